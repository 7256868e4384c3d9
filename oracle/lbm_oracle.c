/*
 * CPU oracle: a plain-C restatement of the reference single-phase D3Q19 MRT
 * solver, Single_phase/LBM_3D_SinglePhase_Solver.py (class
 * LB3D_Solver_Single_Phase, "SP" below), following its unfused four-kernel
 * step  colission -> streaming1 -> Boundary_condition -> streaming3  (SP:477-481)
 * on the reference's own data layout: dense C-order grid, AoS [cell][19]
 * populations, an f/F double buffer, rho and v per cell, and in sparse mode
 * iteration over the active 3x3x3 blocks only (SP:41-44).
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker the HIP path is compared
 * against and the CPU baseline that the benchmark times beside it.  Nothing
 * in the product path links or calls it.
 *
 * Build: REAL=float gives the FP32 restatement (same arithmetic type as the
 * reference), REAL=double the FP64 one.  Built with -ffp-contract=off so each
 * multiply and add rounds separately, in the source order of the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif
#ifndef PFX
#define PFX orc64_
#endif
#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC(name) ORC_CAT(PFX, name)

#define Q 19

/* Direction set, SP:183-187 (e) and SP:189-193 (e_f). */
static const int E[Q][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
    {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
    {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
/* Opposite directions, SP:85. */
static const int LR[Q] __attribute__((unused)) = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

/* Moment matrix M, SP:64-82 (integer entries, same row order). */
static const int MI[Q][Q] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {-1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, -2, -2, -2, -2, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, -2, 2, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
    {0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
    {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
    {0, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
    {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
    {0, 0, 0, -1, -1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1}};

typedef struct {
  int nx, ny, nz;
  int sparse;            /* sparse_storage flag, SP:11,36-44 */
  int px, py, pz;        /* field extent: n, or 3*(n//3+1) when sparse */
  int64_t ncell;         /* px*py*pz */
  int8_t *solid;         /* (nx,ny,nz), SP:57 */
  REAL *f, *F;           /* [ncell][19], SP:32-33 / 37-38 */
  REAL *rho, *v;         /* [ncell], [ncell][3], SP:34-35 / 39-40 */
  uint8_t *active;       /* per 3^3 block (sparse) */
  int nbx, nby, nbz;
  int64_t *act_list;     /* active block ids */
  int64_t n_act;
  /* baked parameters (frozen at init_simulation, SP:118-149) */
  REAL M[Q][Q], invM[Q][Q], S[Q], w[Q], ef[Q][3];
  REAL fx, fy, fz;
  int force_flag;
  int convention;        /* 0 = SP (tau=niu/3+1/2, Guo /3 /9), 1 = documented (tau=3niu+1/2, Guo x1),
                            2 = textbook (tau=3niu+1/2, Guo x3 x9) */
  REAL ga, gb;           /* Guo scalings */
  int bc_mode[6];        /* x-left, x-right, y-left, y-right, z-left, z-right */
  REAL rho_bc[6];
  REAL u_bc[6][3];
  int nthreads;
} ORC(t);

static inline int64_t idx(const ORC(t) * o, int i, int j, int k) {
  return ((int64_t)i * o->py + j) * o->pz + k;
}
static inline int64_t sidx(const ORC(t) * o, int i, int j, int k) {
  return ((int64_t)i * o->ny + j) * o->nz + k;
}

/* np.linalg.inv(M_np) computed in double (SP:83), then stored as REAL (SP:110). */
static void invert_M(double out[Q][Q]) {
  double a[Q][2 * Q];
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < 2 * Q; ++c) a[r][c] = c < Q ? (double)MI[r][c] : (c - Q == r ? 1.0 : 0.0);
  for (int c = 0; c < Q; ++c) {
    int p = c;
    for (int r = c + 1; r < Q; ++r)
      if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
    if (p != c)
      for (int k = 0; k < 2 * Q; ++k) {
        double t = a[c][k];
        a[c][k] = a[p][k];
        a[p][k] = t;
      }
    double d = a[c][c];
    for (int k = 0; k < 2 * Q; ++k) a[c][k] /= d;
    for (int r = 0; r < Q; ++r) {
      if (r == c) continue;
      double m = a[r][c];
      if (m == 0.0) continue;
      for (int k = 0; k < 2 * Q; ++k) a[r][k] -= m * a[c][k];
    }
  }
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < Q; ++c) out[r][c] = a[r][Q + c];
}

ORC(t) * ORC(create)(int nx, int ny, int nz, int sparse) {
  ORC(t) *o = (ORC(t) *)calloc(1, sizeof(ORC(t)));
  o->nx = nx; o->ny = ny; o->nz = nz; o->sparse = sparse;
  if (sparse) { /* pointer((n//3+1)^3) -> dense(3^3), SP:41-44 */
    o->px = 3 * (nx / 3 + 1); o->py = 3 * (ny / 3 + 1); o->pz = 3 * (nz / 3 + 1);
  } else {
    o->px = nx; o->py = ny; o->pz = nz;
  }
  o->nbx = (o->px + 2) / 3; o->nby = (o->py + 2) / 3; o->nbz = (o->pz + 2) / 3;
  o->ncell = (int64_t)o->px * o->py * o->pz;
  o->solid = (int8_t *)calloc((size_t)nx * ny * nz, 1);
  o->f = (REAL *)calloc((size_t)o->ncell * Q, sizeof(REAL));
  o->F = (REAL *)calloc((size_t)o->ncell * Q, sizeof(REAL));
  o->rho = (REAL *)calloc((size_t)o->ncell, sizeof(REAL));
  o->v = (REAL *)calloc((size_t)o->ncell * 3, sizeof(REAL));
  o->active = (uint8_t *)calloc((size_t)o->nbx * o->nby * o->nbz, 1);
  o->act_list = NULL;
  o->nthreads = 1;
  for (int s = 0; s < 6; ++s) o->rho_bc[s] = (REAL)1.0;
  return o;
}

void ORC(destroy)(ORC(t) * o) {
  if (!o) return;
  free(o->solid); free(o->f); free(o->F); free(o->rho); free(o->v);
  free(o->active); free(o->act_list); free(o);
}

void ORC(set_threads)(ORC(t) * o, int n) { o->nthreads = n > 0 ? n : 1; }

/* solid.from_numpy (SP:177): any value > 0 is solid (SP:175). */
void ORC(set_solid)(ORC(t) * o, const int8_t *solid) {
  memcpy(o->solid, solid, (size_t)o->nx * o->ny * o->nz);
}

/* Parameters that SP keeps as Python attributes (SP:17-28, setters SP:405-458). */
void ORC(set_params)(ORC(t) * o, double niu, int convention, double fx, double fy, double fz,
                     const int32_t *bc_mode, const double *rho_bc, const double *u_bc) {
  o->convention = convention;
  o->fx = (REAL)fx; o->fy = (REAL)fy; o->fz = (REAL)fz;
  /* init_simulation: tau, s_v, s_other (SP:126-129), evaluated in Python doubles. */
  double tau = convention == 0 ? niu / 3.0 + 0.5 : 3.0 * niu + 0.5;
  double s_v = 1.0 / tau;
  double s_o = 8.0 * (2.0 - s_v) / (8.0 - s_v);
  const double Sd[Q] = {0, s_v, s_v, 0, s_o, 0, s_o, 0, s_o, s_v, s_v, s_v, s_v, s_v, s_v, s_v, s_o, s_o, s_o};
  for (int s = 0; s < Q; ++s) o->S[s] = (REAL)Sd[s]; /* S_dig, SP:131 */
  o->ga = convention == 0 ? (REAL)(1.0 / 3.0) : (convention == 1 ? (REAL)1.0 : (REAL)3.0);
  o->gb = convention == 0 ? (REAL)(1.0 / 9.0) : (convention == 1 ? (REAL)1.0 : (REAL)9.0);
  /* force_flag, SP:137-140 */
  o->force_flag = (fabs(fx) > 0 || fabs(fy) > 0 || fabs(fz) > 0) ? 1 : 0;
  for (int s = 0; s < 6; ++s) {
    o->bc_mode[s] = bc_mode[s];
    o->rho_bc[s] = (REAL)rho_bc[s];
    for (int c = 0; c < 3; ++c) o->u_bc[s][c] = (REAL)u_bc[3 * s + c];
  }
}

/* What SP's kernels capture when they are first compiled, at the first step():
 * the body force of cal_local_force (SP:217-220) and the face modes and
 * densities that Boundary_condition tests with ti.static (SP:274-365).  S,
 * force_flag and the face velocity lists stay as init_simulation froze them
 * (SP:119-140). */
void ORC(set_compile_params)(ORC(t) * o, double fx, double fy, double fz, const int32_t *bc_mode,
                             const double *rho_bc) {
  o->fx = (REAL)fx; o->fy = (REAL)fy; o->fz = (REAL)fz;
  for (int s = 0; s < 6; ++s) {
    o->bc_mode[s] = bc_mode[s];
    o->rho_bc[s] = (REAL)rho_bc[s];
  }
}

/* static_init (SP:180-197) and the M / inv_M fields (SP:61-110). */
static void static_init(ORC(t) * o) {
  double inv[Q][Q];
  invert_M(inv);
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < Q; ++c) {
      o->M[r][c] = (REAL)MI[r][c];
      o->invM[r][c] = (REAL)inv[r][c];
    }
  for (int s = 0; s < Q; ++s) {
    for (int c = 0; c < 3; ++c) o->ef[s][c] = (REAL)E[s][c];
    o->w[s] = s == 0 ? (REAL)(1.0 / 3.0) : (s < 7 ? (REAL)(1.0 / 18.0) : (REAL)(1.0 / 36.0));
  }
}

/* feq, SP:152-158 */
static inline REAL feq(const ORC(t) * o, int k, REAL rho, const REAL u[3]) {
  REAL eu = (REAL)E[k][0] * u[0] + (REAL)E[k][1] * u[1] + (REAL)E[k][2] * u[2];
  REAL uv = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  return o->w[k] * rho * ((REAL)1.0 + (REAL)3.0 * eu + (REAL)4.5 * eu * eu - (REAL)1.5 * uv);
}

static inline int is_fluid_cell(const ORC(t) * o, int i, int j, int k) {
  return i < o->nx && j < o->ny && k < o->nz && o->solid[sidx(o, i, j, k)] == 0;
}

/* init kernel, SP:160-170.  In sparse mode only fluid cells are written, which
 * activates their 3^3 blocks (SP:164). */
static void init_fields(ORC(t) * o) {
  const REAL u0[3] = {0, 0, 0};
  REAL f0[Q];
  for (int s = 0; s < Q; ++s) f0[s] = feq(o, s, (REAL)1.0, u0);
  memset(o->active, 0, (size_t)o->nbx * o->nby * o->nbz);
#pragma omp parallel for schedule(static) num_threads(o->nthreads)
  for (int i = 0; i < o->nx; ++i)
    for (int j = 0; j < o->ny; ++j)
      for (int k = 0; k < o->nz; ++k) {
        if (o->sparse && o->solid[sidx(o, i, j, k)] != 0) continue;
        int64_t c = idx(o, i, j, k);
        o->rho[c] = (REAL)1.0;
        o->v[3 * c] = o->v[3 * c + 1] = o->v[3 * c + 2] = 0;
        for (int s = 0; s < Q; ++s) {
          o->f[c * Q + s] = f0[s];
          o->F[c * Q + s] = f0[s];
        }
        o->active[((int64_t)(i / 3) * o->nby + j / 3) * o->nbz + k / 3] = 1; /* same value from any thread */
      }
  free(o->act_list);
  int64_t nb = (int64_t)o->nbx * o->nby * o->nbz;
  o->act_list = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1));
  o->n_act = 0;
  for (int64_t b = 0; b < nb; ++b)
    if (!o->sparse || o->active[b]) o->act_list[o->n_act++] = b;
}

void ORC(init_simulation)(ORC(t) * o) {
  static_init(o);
  init_fields(o);
}

/* The struct-for domain of SP's kernels: every cell (dense) or every cell of
 * an active 3^3 block (sparse).  Blocks are visited in parallel; the body
 * callbacks only touch what the reference's per-cell body touches. */
#define FOR_DOMAIN(o, ...)                                                              \
  do {                                                                                   \
    _Pragma("omp parallel for schedule(static) num_threads(o->nthreads)")                \
    for (int64_t ab = 0; ab < o->n_act; ++ab) {                                          \
      int64_t b = o->act_list[ab];                                                       \
      int bi = (int)(b / ((int64_t)o->nby * o->nbz));                                    \
      int bj = (int)((b / o->nbz) % o->nby);                                             \
      int bk = (int)(b % o->nbz);                                                        \
      for (int di = 0; di < 3; ++di)                                                     \
        for (int dj = 0; dj < 3; ++dj)                                                   \
          for (int dk = 0; dk < 3; ++dk) {                                               \
            int i = 3 * bi + di, j = 3 * bj + dj, k = 3 * bk + dk;                       \
            if (i >= o->px || j >= o->py || k >= o->pz) continue;                        \
            __VA_ARGS__                                                                  \
          }                                                                              \
    }                                                                                    \
  } while (0)

/* colission, SP:222-241 */
static void colission(ORC(t) * o) {
  FOR_DOMAIN(o, {
    if (is_fluid_cell(o, i, j, k)) {
      int64_t c = idx(o, i, j, k);
      const REAL *Fc = o->F + c * Q;
      REAL m[Q], meq[Q];
      for (int s = 0; s < Q; ++s) { /* M[None] @ F[i,j,k]  (SP:226) */
        REAL acc = o->M[s][0] * Fc[0];
        for (int l = 1; l < Q; ++l) acc = acc + o->M[s][l] * Fc[l];
        m[s] = acc;
      }
      const REAL *u = o->v + 3 * c;
      /* meq_vec, SP:209-215 */
      for (int s = 0; s < Q; ++s) meq[s] = 0;
      meq[0] = o->rho[c]; meq[3] = u[0]; meq[5] = u[1]; meq[7] = u[2];
      meq[1] = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      meq[9] = (REAL)2 * u[0] * u[0] - u[1] * u[1] - u[2] * u[2];
      meq[11] = u[1] * u[1] - u[2] * u[2];
      meq[13] = u[0] * u[1]; meq[14] = u[1] * u[2]; meq[15] = u[0] * u[2];
      for (int s = 0; s < Q; ++s) m[s] = m[s] - o->S[s] * (m[s] - meq[s]); /* SP:228 */
      if (o->force_flag) { /* Guo forcing, SP:230-238 */
        const REAL fv[3] = {o->fx, o->fy, o->fz};
        for (int s = 0; s < Q; ++s) {
          REAL f_guo = 0;
          for (int l = 0; l < Q; ++l) {
            REAL emu = (o->ef[l][0] - u[0]) * fv[0] + (o->ef[l][1] - u[1]) * fv[1] +
                       (o->ef[l][2] - u[2]) * fv[2];
            REAL eu = o->ef[l][0] * u[0] + o->ef[l][1] * u[1] + o->ef[l][2] * u[2];
            REAL ef = o->ef[l][0] * fv[0] + o->ef[l][1] * fv[1] + o->ef[l][2] * fv[2];
            REAL term;
            if (o->convention == 0)
              term = emu / (REAL)3.0 + (eu * ef) / (REAL)9.0;
            else
              term = emu * o->ga + (eu * ef) * o->gb;
            f_guo = f_guo + o->w[l] * term * o->M[s][l];
          }
          m[s] = m[s] + ((REAL)1 - (REAL)0.5 * o->S[s]) * f_guo;
        }
      }
      REAL *fc = o->f + c * Q; /* f = 0; f += inv_M @ m  (SP:240-241) */
      for (int s = 0; s < Q; ++s) {
        REAL acc = o->invM[s][0] * m[0];
        for (int l = 1; l < Q; ++l) acc = acc + o->invM[s][l] * m[l];
        fc[s] = (REAL)0 + acc;
      }
    }
  });
}

/* periodic_index, SP:247-257 */
static inline void periodic_index(const ORC(t) * o, int *a) {
  if (a[0] < 0) a[0] = o->nx - 1;
  if (a[0] > o->nx - 1) a[0] = 0;
  if (a[1] < 0) a[1] = o->ny - 1;
  if (a[1] > o->ny - 1) a[1] = 0;
  if (a[2] < 0) a[2] = o->nz - 1;
  if (a[2] > o->nz - 1) a[2] = 0;
}

/* streaming1, SP:259-269: push, or halfway bounce-back into the opposite slot. */
static void streaming1(ORC(t) * o) {
  FOR_DOMAIN(o, {
    if (is_fluid_cell(o, i, j, k)) {
      int64_t c = idx(o, i, j, k);
      for (int s = 0; s < Q; ++s) {
        int ip[3] = {i + E[s][0], j + E[s][1], k + E[s][2]};
        periodic_index(o, ip);
        if (o->solid[sidx(o, ip[0], ip[1], ip[2])] == 0)
          o->F[idx(o, ip[0], ip[1], ip[2]) * Q + s] = o->f[c * Q + s];
        else
          o->F[c * Q + LR[s]] = o->f[c * Q + s];
      }
    }
  });
}

/* One face of Boundary_condition (SP:272-370).  face: 0 x-left, 1 x-right,
 * 2 y-left, 3 y-right, 4 z-left, 5 z-right.  (a,b) runs over the face plane. */
static void face_bc(ORC(t) * o, int face) {
  int axis = face / 2, right = face % 2;
  int n[3] = {o->nx, o->ny, o->nz};
  int pos = right ? n[axis] - 1 : 0;
  int inw = right ? n[axis] - 2 : 1;
  int ax1 = axis == 0 ? 1 : 0, ax2 = axis == 2 ? 1 : 2;
  int mode = o->bc_mode[face];
  if (mode != 1 && mode != 2) return; /* only codes 1 and 2 are acted on */
  if (n[axis] < 2) return;            /* no inward neighbour exists (callers reject this) */
  for (int a = 0; a < n[ax1]; ++a)
    for (int b = 0; b < n[ax2]; ++b) {
      int p[3], q[3];
      p[axis] = pos; p[ax1] = a; p[ax2] = b;
      q[axis] = inw; q[ax1] = a; q[ax2] = b;
      if (o->solid[sidx(o, p[0], p[1], p[2])] != 0) continue;
      int64_t c = idx(o, p[0], p[1], p[2]);
      if (mode == 1) { /* fixed pressure: SP:274-281 (x-left) and the five like it */
        int64_t cq = idx(o, q[0], q[1], q[2]);
        const REAL *u = o->solid[sidx(o, q[0], q[1], q[2])] > 0 ? o->v + 3 * cq : o->v + 3 * c;
        REAL uu[3] = {u[0], u[1], u[2]};
        for (int s = 0; s < Q; ++s) o->F[c * Q + s] = feq(o, s, o->rho_bc[face], uu);
      } else { /* fixed velocity: SP:283-288 (x-left) and the five like it */
        for (int s = 0; s < Q; ++s) o->F[c * Q + s] = feq(o, s, (REAL)1.0, o->u_bc[face]);
      }
    }
}

static void boundary_condition(ORC(t) * o) {
  for (int face = 0; face < 6; ++face) face_bc(o, face); /* source order, SP:274-370 */
}

/* streaming3, SP:372-392 */
static void streaming3(ORC(t) * o) {
  FOR_DOMAIN(o, {
    int64_t c = idx(o, i, j, k);
    if (is_fluid_cell(o, i, j, k)) {
      REAL *fc = o->f + c * Q;
      const REAL *Fc = o->F + c * Q;
      for (int s = 0; s < Q; ++s) fc[s] = Fc[s];
      REAL sum = fc[0];
      for (int s = 1; s < Q; ++s) sum = sum + fc[s];
      REAL r = (REAL)0 + sum;
      REAL vx = 0, vy = 0, vz = 0;
      for (int s = 0; s < Q; ++s) {
        vx = vx + o->ef[s][0] * fc[s];
        vy = vy + o->ef[s][1] * fc[s];
        vz = vz + o->ef[s][2] * fc[s];
      }
      vx = vx / r; vy = vy / r; vz = vz / r;
      vx = vx + (o->fx / (REAL)2) / r;
      vy = vy + (o->fy / (REAL)2) / r;
      vz = vz + (o->fz / (REAL)2) / r;
      o->rho[c] = r;
      o->v[3 * c] = vx; o->v[3 * c + 1] = vy; o->v[3 * c + 2] = vz;
    } else {
      o->rho[c] = (REAL)1.0;
      o->v[3 * c] = o->v[3 * c + 1] = o->v[3 * c + 2] = 0;
    }
  });
}

/* step, SP:477-481 */
void ORC(step)(ORC(t) * o, int nsteps) {
  for (int t = 0; t < nsteps; ++t) {
    colission(o);
    streaming1(o);
    boundary_condition(o);
    streaming3(o);
  }
}

/* get_max_v / cal_max_v, SP:394-402: max |v| over every cell of the domain. */
double ORC(get_max_v)(ORC(t) * o) {
  REAL mx = (REAL)-1e10;
  for (int64_t ab = 0; ab < o->n_act; ++ab) {
    int64_t b = o->act_list[ab];
    int bi = (int)(b / ((int64_t)o->nby * o->nbz)), bj = (int)((b / o->nbz) % o->nby), bk = (int)(b % o->nbz);
    for (int di = 0; di < 3; ++di)
      for (int dj = 0; dj < 3; ++dj)
        for (int dk = 0; dk < 3; ++dk) {
          int i = 3 * bi + di, j = 3 * bj + dj, k = 3 * bk + dk;
          if (i >= o->px || j >= o->py || k >= o->pz) continue;
          const REAL *u = o->v + 3 * idx(o, i, j, k);
          REAL nrm = (REAL)sqrt((double)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
          if (nrm > mx) mx = nrm;
        }
  }
  return (double)mx;
}

/* Field read-back (to_numpy, C-order, the padded extent in sparse mode). */
void ORC(get_shape)(ORC(t) * o, int32_t *shape) { shape[0] = o->px; shape[1] = o->py; shape[2] = o->pz; }
void ORC(get_rho)(ORC(t) * o, double *out) {
  for (int64_t c = 0; c < o->ncell; ++c) out[c] = (double)o->rho[c];
}
void ORC(get_v)(ORC(t) * o, double *out) {
  for (int64_t c = 0; c < 3 * o->ncell; ++c) out[c] = (double)o->v[c];
}
void ORC(get_F)(ORC(t) * o, double *out) {
  for (int64_t c = 0; c < Q * o->ncell; ++c) out[c] = (double)o->F[c];
}
void ORC(set_F)(ORC(t) * o, const double *in) { /* test hook: arbitrary population state */
  for (int64_t c = 0; c < Q * o->ncell; ++c) o->F[c] = (REAL)in[c];
}
void ORC(set_macros)(ORC(t) * o, const double *rho, const double *v) { /* test hook */
  for (int64_t c = 0; c < o->ncell; ++c) o->rho[c] = (REAL)rho[c];
  for (int64_t c = 0; c < 3 * o->ncell; ++c) o->v[c] = (REAL)v[c];
}
void ORC(get_invM)(ORC(t) * o, double *out) {
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < Q; ++c) out[r * Q + c] = (double)o->invM[r][c];
}
void ORC(get_M)(ORC(t) * o, double *out) {
  for (int r = 0; r < Q; ++r)
    for (int c = 0; c < Q; ++c) out[r * Q + c] = (double)o->M[r][c];
}
int64_t ORC(n_active_blocks)(ORC(t) * o) { return o->n_act; }
int ORC(real_bytes)(void) { return (int)sizeof(REAL); }
