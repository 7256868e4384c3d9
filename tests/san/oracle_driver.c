/*
 * Sanitizer driver for the C oracle (oracle/lbm_oracle.c): runs the
 * restatement of SP's four-kernel step over small grids that reach every
 * branch (dense and sparse storage, extents of 1 to 3 along an axis, solid
 * faces, pressure and velocity faces on every side, all three force
 * conventions, forced and unforced) and reads every field back.  Built with
 * ASan + UBSan by tests/san/Makefile; any report aborts with a non-zero exit.
 * Prints one line per case and "ok" at the end.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC(name) ORC_CAT(PFX, name)

typedef struct ORC(t) ORC(t);
ORC(t) * ORC(create)(int nx, int ny, int nz, int sparse);
void ORC(destroy)(ORC(t) * o);
void ORC(set_threads)(ORC(t) * o, int n);
void ORC(set_solid)(ORC(t) * o, const int8_t *solid);
void ORC(set_params)(ORC(t) * o, double niu, int convention, double fx, double fy, double fz,
                     const int32_t *bc_mode, const double *rho_bc, const double *u_bc);
void ORC(set_compile_params)(ORC(t) * o, double fx, double fy, double fz, const int32_t *bc_mode,
                             const double *rho_bc);
void ORC(init_simulation)(ORC(t) * o);
void ORC(step)(ORC(t) * o, int nsteps);
double ORC(get_max_v)(ORC(t) * o);
void ORC(get_shape)(ORC(t) * o, int32_t *shape);
void ORC(get_rho)(ORC(t) * o, double *out);
void ORC(get_v)(ORC(t) * o, double *out);
void ORC(get_F)(ORC(t) * o, double *out);
void ORC(set_F)(ORC(t) * o, const double *in);
void ORC(set_macros)(ORC(t) * o, const double *rho, const double *v);
void ORC(get_invM)(ORC(t) * o, double *out);
void ORC(get_M)(ORC(t) * o, double *out);
int64_t ORC(n_active_blocks)(ORC(t) * o);

static uint32_t rng = 12345u;
static double urand(void) {
  rng ^= rng << 13;
  rng ^= rng >> 17;
  rng ^= rng << 5;
  return (double)rng / 4294967296.0;
}

static int run_case(int nx, int ny, int nz, int sparse, int conv, int forced, const int32_t mode[6], double porosity,
                    int threads) {
  ORC(t) *o = ORC(create)(nx, ny, nz, sparse);
  ORC(set_threads)(o, threads);
  const size_t n = (size_t)nx * ny * nz;
  int8_t *solid = (int8_t *)malloc(n);
  for (size_t c = 0; c < n; ++c) solid[c] = urand() < porosity ? 0 : (urand() < 0.5 ? 1 : 2);
  ORC(set_solid)(o, solid);
  const double rho_bc[6] = {1.004, 0.997, 1.002, 0.999, 1.001, 0.998};
  const double u_bc[18] = {0.01, 0, 0, 0.02, 0, 0, 0, 0.01, 0, 0, -0.01, 0, 0, 0, 0.02, 0.03, 0, 0};
  ORC(set_params)(o, 0.1, conv, forced ? 1e-4 : 0.0, forced ? -2e-5 : 0.0, 0.0, mode, rho_bc, u_bc);
  ORC(init_simulation)(o);
  ORC(set_compile_params)(o, forced ? 1e-4 : 0.0, forced ? -2e-5 : 0.0, 0.0, mode, rho_bc);
  ORC(step)(o, 7);
  int32_t sh[3];
  ORC(get_shape)(o, sh);
  const size_t cells = (size_t)sh[0] * sh[1] * sh[2];
  double *rho = (double *)malloc(cells * sizeof(double));
  double *v = (double *)malloc(3 * cells * sizeof(double));
  double *F = (double *)malloc(19 * cells * sizeof(double));
  ORC(get_rho)(o, rho);
  ORC(get_v)(o, v);
  ORC(get_F)(o, F);
  ORC(set_F)(o, F);
  ORC(set_macros)(o, rho, v);
  ORC(step)(o, 2);
  const double mv = ORC(get_max_v)(o);
  double sum = 0;
  for (size_t c = 0; c < cells; ++c) sum += rho[c];
  double M[361], invM[361];
  ORC(get_M)(o, M);
  ORC(get_invM)(o, invM);
  printf("%dx%dx%d sparse %d conv %d forced %d modes %d%d%d%d%d%d blocks %lld sum_rho %.6f max_v %.3e\n", nx, ny, nz,
         sparse, conv, forced, mode[0], mode[1], mode[2], mode[3], mode[4], mode[5],
         (long long)ORC(n_active_blocks)(o), sum, mv);
  const int bad = !isfinite(sum) || !isfinite(mv);
  free(rho);
  free(v);
  free(F);
  free(solid);
  ORC(destroy)(o);
  return bad;
}

int main(void) {
  const int32_t periodic[6] = {0, 0, 0, 0, 0, 0};
  const int32_t pressure_x[6] = {1, 1, 0, 0, 0, 0};
  const int32_t mixed[6] = {1, 2, 2, 1, 1, 2};
  const int32_t all_vel[6] = {2, 2, 2, 2, 2, 2};
  const int32_t all_rho[6] = {1, 1, 1, 1, 1, 1};
  const int32_t *modes[5] = {periodic, pressure_x, mixed, all_vel, all_rho};
  const int shapes[][3] = {{5, 4, 3}, {2, 3, 7}, {3, 3, 3}, {7, 2, 2}, {4, 5, 6}, {1, 4, 5}};
  int bad = 0;
  for (int s = 0; s < 6; ++s)
    for (int m = 0; m < 5; ++m) {
      const int *e = shapes[s];
      if (e[0] < 2 && m > 0) continue;  // faces need an extent >= 2 along their normal
      for (int sparse = 0; sparse < 2; ++sparse)
        bad |= run_case(e[0], e[1], e[2], sparse, (s + m) % 3, (s + m + sparse) & 1, modes[m], 0.7, 1 + (m & 1));
    }
  bad |= run_case(6, 5, 4, 1, 0, 1, mixed, 0.0, 2);  // no fluid cell at all
  bad |= run_case(6, 5, 4, 0, 2, 1, pressure_x, 1.0, 2);  // all fluid
  if (bad) {
    printf("non-finite result\n");
    return 1;
  }
  printf("ok\n");
  return 0;
}
