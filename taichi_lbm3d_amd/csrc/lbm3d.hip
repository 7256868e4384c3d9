// liblbm3d.so -- single-phase sparse D3Q19 MRT lattice Boltzmann for MI355X (gfx950).
//
// Replaces the Taichi kernels of the reference class LB3D_Solver_Single_Phase
// (Single_phase/LBM_3D_SinglePhase_Solver.py, "SP"): its four-kernel step
// colission -> streaming1 -> Boundary_condition -> streaming3 (SP:477-481)
// becomes one fused kernel over a compacted fluid-cell list, in place with the
// AA access pattern, with the MRT transform held in registers.  The C ABI is
// declared in include/lbm3d.h.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "lbm3d.h"

namespace {

constexpr int kQ = 19;
constexpr int kBlock = 256;

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, const char *what) {
  snprintf(g_err, sizeof(g_err), fmt, what);
  return code;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return LBM3D_EHIP;
  }
  return LBM3D_OK;
}

// ---------------------------------------------------------------------------
// Lattice constants.  Direction order of SP:183-187; opposite map of SP:85.
// Directions 2p-1 and 2p (p = 1..9) are an opposite pair.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__host__ __device__ constexpr int ex(int q) {
  constexpr int t[kQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[q];
}
__host__ __device__ constexpr int ey(int q) {
  constexpr int t[kQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[q];
}
__host__ __device__ constexpr int ez(int q) {
  constexpr int t[kQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[q];
}
// weights, SP:195-197
__host__ __device__ constexpr float wq(int q) {
  return q == 0 ? 1.0f / 3.0f : (q < 7 ? 1.0f / 18.0f : 1.0f / 36.0f);
}

// ---------------------------------------------------------------------------
// Storage conversion: populations live in HBM as g = f - w in T, compute in FP32.
// ---------------------------------------------------------------------------
template <typename T> struct St;
template <> struct St<float> {
  static __device__ __forceinline__ float ld(const float *p) { return *p; }
  static __device__ __forceinline__ void st(float *p, float v) { *p = v; }
};
template <> struct St<__half> {
  static __device__ __forceinline__ float ld(const __half *p) { return __half2float(*p); }
  static __device__ __forceinline__ void st(__half *p, float v) { *p = __float2half_rn(v); }
};
template <> struct St<__hip_bfloat16> {
  static __device__ __forceinline__ float ld(const __hip_bfloat16 *p) { return __bfloat162float(*p); }
  static __device__ __forceinline__ void st(__hip_bfloat16 *p, float v) { *p = __float2bfloat16(v); }
};

// ---------------------------------------------------------------------------
// MRT collision on shifted populations g (g_q = f_q - w_q), in registers.
//
// The reference computes f = M^-1 [ m - S (m - meq) + (1 - S/2) G ] with
// m = M F (SP:226-241).  Here:
//  * M is applied through opposite-pair sums P and differences D (even rows
//    of M see only P, odd rows only D), every row being a polynomial in e:
//    1, |e|^2-1, 3|e|^4-6|e|^2+1, e_x, (3|e|^2-5)e_x, ..., 3e_x^2-|e|^2, ...
//  * M^-1 = M^T (M M^T)^-1 with M M^T block diagonal ({0,1,2}, {3,4}, {5,6},
//    {7,8}, {9,10}, {11,12} and a diagonal), so the inverse is a few 2x2/3x3
//    products followed by the integer M^T, again split into pair halves.
//  * meq is the incompressible form of meq_vec (SP:209-215); it depends on u
//    only for every moment but m0, whose rate S_0 is 0.  Because M w = e_0,
//    shifting by w changes m0 alone (m0 = rho - 1), so the whole collision
//    runs on g.
//  * G_s = sum_l w_l [a (e_l - u).F + b (e_l.u)(e_l.F)] M[s,l] (SP:231-236)
//    has the closed form below (nonzero only for s = 0,1,3,5,7,9,11,13,14,15).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mrt_collide(float g[kQ], float ux, float uy, float uz,
                                            const lbm3d_params &p) {
  const float sv = p.s_v, so = p.s_o;
  float P[10], D[10];
#pragma unroll
  for (int k = 1; k <= 9; ++k) {
    P[k] = g[2 * k - 1] + g[2 * k];
    D[k] = g[2 * k - 1] - g[2 * k];
  }
  const float g0 = g[0];
  const float s13 = P[1] + P[2] + P[3];
  const float s49 = (P[4] + P[5]) + (P[6] + P[7]) + (P[8] + P[9]);
  const float p4567 = (P[4] + P[5]) + (P[6] + P[7]);
  const float p89 = P[8] + P[9];

  float m0 = g0 + s13 + s49;  // rho - 1
  float m1 = s49 - g0;
  float m2 = g0 - 2.0f * s13 + s49;
  float m9 = 2.0f * P[1] - P[2] - P[3] + p4567 - 2.0f * p89;
  float m10 = -2.0f * P[1] + P[2] + P[3] + p4567 - 2.0f * p89;
  float m11 = P[2] - P[3] + (P[4] + P[5]) - (P[6] + P[7]);
  float m12 = P[3] - P[2] + (P[4] + P[5]) - (P[6] + P[7]);
  float m13 = P[4] - P[5];
  float m14 = P[8] - P[9];
  float m15 = P[6] - P[7];
  const float dx4 = (D[4] + D[5]) + (D[6] + D[7]);
  const float dy4 = (D[4] - D[5]) + (D[8] + D[9]);
  const float dz4 = (D[6] - D[7]) + (D[8] - D[9]);
  float m3 = D[1] + dx4;
  float m4 = dx4 - 2.0f * D[1];
  float m5 = D[2] + dy4;
  float m6 = dy4 - 2.0f * D[2];
  float m7 = D[3] + dz4;
  float m8 = dz4 - 2.0f * D[3];
  float m16 = (D[4] + D[5]) - (D[6] + D[7]);
  float m17 = (D[5] - D[4]) + (D[8] + D[9]);
  float m18 = (D[6] - D[7]) - (D[8] - D[9]);

  // relaxation towards meq (SP:227-228)
  const float uxx = ux * ux, uyy = uy * uy, uzz = uz * uz;
  m1 -= sv * (m1 - (uxx + uyy + uzz));
  m2 -= sv * m2;
  m4 -= so * m4;
  m6 -= so * m6;
  m8 -= so * m8;
  m9 -= sv * (m9 - (2.0f * uxx - uyy - uzz));
  m10 -= sv * m10;
  m11 -= sv * (m11 - (uyy - uzz));
  m12 -= sv * m12;
  m13 -= sv * (m13 - ux * uy);
  m14 -= sv * (m14 - uy * uz);
  m15 -= sv * (m15 - ux * uz);
  m16 -= so * m16;
  m17 -= so * m17;
  m18 -= so * m18;

  // forcing in moment space (SP:230-238), closed form of M g_guo
  if (p.force_on) {
    const float a = p.guo_a, b = p.guo_b;
    const float Fx = p.fx, Fy = p.fy, Fz = p.fz;
    const float uF = ux * Fx + uy * Fy + uz * Fz;
    const float hv = 1.0f - 0.5f * sv;
    const float b29 = b * (2.0f / 9.0f), b9 = b * (1.0f / 9.0f);
    m0 += -(3.0f * a - b) * (1.0f / 3.0f) * uF;
    m1 += hv * b29 * uF;
    m3 += a * (1.0f / 3.0f) * Fx;
    m5 += a * (1.0f / 3.0f) * Fy;
    m7 += a * (1.0f / 3.0f) * Fz;
    m9 += hv * b29 * (2.0f * Fx * ux - Fy * uy - Fz * uz);
    m11 += hv * b29 * (Fy * uy - Fz * uz);
    m13 += hv * b9 * (Fx * uy + Fy * ux);
    m14 += hv * b9 * (Fy * uz + Fz * uy);
    m15 += hv * b9 * (Fx * uz + Fz * ux);
  }

  // y = (M M^T)^-1 m'
  const float y0 = (5.0f / 36.0f) * m0 - (11.0f / 72.0f) * m1 + (1.0f / 24.0f) * m2;
  const float y1 = -(11.0f / 72.0f) * m0 + (13.0f / 48.0f) * m1 - (11.0f / 144.0f) * m2;
  const float y2 = (1.0f / 24.0f) * m0 - (11.0f / 144.0f) * m1 + (7.0f / 144.0f) * m2;
  const float y3 = (1.0f / 9.0f) * m3 - (1.0f / 36.0f) * m4;
  const float y4 = -(1.0f / 36.0f) * m3 + (5.0f / 72.0f) * m4;
  const float y5 = (1.0f / 9.0f) * m5 - (1.0f / 36.0f) * m6;
  const float y6 = -(1.0f / 36.0f) * m5 + (5.0f / 72.0f) * m6;
  const float y7 = (1.0f / 9.0f) * m7 - (1.0f / 36.0f) * m8;
  const float y8 = -(1.0f / 36.0f) * m7 + (5.0f / 72.0f) * m8;
  const float y9 = (1.0f / 32.0f) * m9 - (1.0f / 96.0f) * m10;
  const float y10 = -(1.0f / 96.0f) * m9 + (1.0f / 32.0f) * m10;
  const float y11 = (3.0f / 32.0f) * m11 - (1.0f / 32.0f) * m12;
  const float y12 = -(1.0f / 32.0f) * m11 + (3.0f / 32.0f) * m12;
  const float y13 = 0.25f * m13, y14 = 0.25f * m14, y15 = 0.25f * m15;
  const float y16 = 0.125f * m16, y17 = 0.125f * m17, y18 = 0.125f * m18;

  // g' = M^T y, even half E and odd half O per opposite pair
  const float A = y0 + y1 + y2, B = y9 + y10, C = y11 + y12;
  const float a02 = y0 - 2.0f * y2;
  const float b910 = y10 - y9;
  float E[10], O[10];
  E[1] = a02 - 2.0f * b910;
  E[2] = a02 + b910 + (y11 - y12);
  E[3] = a02 + b910 - (y11 - y12);
  E[4] = A + B + C + y13;
  E[5] = A + B + C - y13;
  E[6] = A + B - C + y15;
  E[7] = A + B - C - y15;
  E[8] = A - 2.0f * B + y14;
  E[9] = A - 2.0f * B - y14;
  const float X = y3 + y4, Y = y5 + y6, Z = y7 + y8;
  O[1] = y3 - 2.0f * y4;
  O[2] = y5 - 2.0f * y6;
  O[3] = y7 - 2.0f * y8;
  O[4] = X + Y + (y16 - y17);
  O[5] = X - Y + (y16 + y17);
  O[6] = X + Z - (y16 - y18);
  O[7] = X - Z - (y16 + y18);
  O[8] = Y + Z + (y17 - y18);
  O[9] = Y - Z + (y17 + y18);
  g[0] = y0 - y1 + y2;
#pragma unroll
  for (int k = 1; k <= 9; ++k) {
    g[2 * k - 1] = E[k] + O[k];
    g[2 * k] = E[k] - O[k];
  }
}

// feq(q, rho, u) - w_q  (SP:152-158, shifted)
__device__ __forceinline__ float feq_shift(int q, float rho, float ux, float uy, float uz, float uu) {
  const float eu = (float)ex(q) * ux + (float)ey(q) * uy + (float)ez(q) * uz;
  return wq(q) * (rho - 1.0f) + wq(q) * rho * (3.0f * eu + 4.5f * eu * eu - 1.5f * uu);
}

// ---------------------------------------------------------------------------
// The read stage shared by the step and macros kernels.
//
// Even t: cell x's populations sit in its own slots g[q][x].
// Odd t : population q of cell x sits in slot opp(q) of the cell at x - e_q,
//         or, when that cell is solid, in slot q of x itself (the bounce-back
//         written by x on the previous step).  Equivalently, for each q the
//         address  A_q = (nbr_q(x) >= 0) ? g[q][nbr_q(x)] : g[opp q][x]
//         holds population opp(q); the step then writes post-collision f_q
//         back to A_q, so every cell reads and writes the same set of slots
//         and cells' sets are disjoint (race free in place).
// ---------------------------------------------------------------------------
template <typename T, bool ODD>
__device__ __forceinline__ void read_cell(const T *__restrict__ g, const int32_t *__restrict__ nbr,
                                          uint32_t x, uint32_t N, float v[kQ], uint32_t off[kQ]) {
  if constexpr (!ODD) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) v[q] = St<T>::ld(g + (size_t)q * N + x);
  } else {
    off[0] = x;
#pragma unroll
    for (int q = 1; q < kQ; ++q) {
      const int32_t c = nbr[(size_t)(q - 1) * N + x];
      off[q] = c >= 0 ? (uint32_t)q * N + (uint32_t)c : (uint32_t)opp(q) * N + x;
    }
    v[0] = St<T>::ld(g + x);
#pragma unroll
    for (int q = 1; q < kQ; ++q) v[opp(q)] = St<T>::ld(g + off[q]);
  }
}

// Boundary_condition (SP:272-370) for a face-owned cell: every population is
// replaced by feq of the owner face.  Pressure (mode 1): rho_bc and the cell's
// previous-step velocity, or 0 when the inward neighbour is solid (its v is 0,
// SP:390-392).  Velocity (mode 2): rho 1 and the face velocity.
__device__ __forceinline__ void face_populations(float v[kQ], int32_t tag, const float *face_u,
                                                 uint32_t fi, uint32_t nf, const lbm3d_params &p) {
  const int face = LBM3D_TAG_FACE(tag);
  const int mode = LBM3D_TAG_MODE(tag);
  float rho, ux, uy, uz;
  if (mode == 1) {
    rho = p.rho_bc[face];
    if (LBM3D_TAG_INWARD_SOLID(tag)) {
      ux = uy = uz = 0.0f;
    } else {
      ux = face_u[fi];
      uy = face_u[nf + fi];
      uz = face_u[2 * nf + fi];
    }
  } else {
    rho = 1.0f;
    ux = p.u_bc[face][0];
    uy = p.u_bc[face][1];
    uz = p.u_bc[face][2];
  }
  const float uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int q = 0; q < kQ; ++q) v[q] = feq_shift(q, rho, ux, uy, uz, uu);
}

// rho and v of streaming3 (SP:380-388) from shifted populations.  `first`
// drops the half-force shift: the state left by init carries v = u0 exactly.
__device__ __forceinline__ void macros_of(const float v[kQ], const lbm3d_params &p, bool first,
                                          float &rho, float &ux, float &uy, float &uz) {
  float dr = v[0];
  float jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
  for (int k = 1; k <= 9; ++k) {
    const float P = v[2 * k - 1] + v[2 * k];
    const float D = v[2 * k - 1] - v[2 * k];
    dr += P;
    jx += (float)ex(2 * k - 1) * D;
    jy += (float)ey(2 * k - 1) * D;
    jz += (float)ez(2 * k - 1) * D;
  }
  rho = 1.0f + dr;
  const float inv = 1.0f / rho;
  if (first) {
    ux = jx * inv;
    uy = jy * inv;
    uz = jz * inv;
  } else {
    ux = (jx + 0.5f * p.fx) * inv;
    uy = (jy + 0.5f * p.fy) * inv;
    uz = (jz + 0.5f * p.fz) * inv;
  }
}

template <typename T, bool ODD>
__global__ void __launch_bounds__(kBlock) k_step(const lbm3d_params p, T *__restrict__ g,
                                                 const int32_t *__restrict__ nbr,
                                                 const int32_t *__restrict__ face_tag,
                                                 float *__restrict__ face_u, int first) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)p.n_cells;
  if (x >= N) return;
  float v[kQ];
  uint32_t off[kQ];
  read_cell<T, ODD>(g, nbr, x, N, v, off);

  const bool face = x >= (uint32_t)p.n_interior;
  int32_t tag = 0;
  uint32_t fi = 0;
  const uint32_t nf = N - (uint32_t)p.n_interior;
  if (face) {
    fi = x - (uint32_t)p.n_interior;
    tag = face_tag[fi];
    if (!first) face_populations(v, tag, face_u, fi, nf, p);
  }
  float rho, ux, uy, uz;
  macros_of(v, p, first != 0, rho, ux, uy, uz);
  if (face && LBM3D_TAG_MODE(tag) == 1 && !LBM3D_TAG_INWARD_SOLID(tag)) {
    // the pressure face of the next step uses this step's v (SP:281)
    face_u[fi] = ux;
    face_u[nf + fi] = uy;
    face_u[2 * nf + fi] = uz;
  }
  mrt_collide(v, ux, uy, uz, p);

  if constexpr (!ODD) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) St<T>::st(g + (size_t)opp(q) * N + x, v[q]);
  } else {
    St<T>::st(g + x, v[0]);
#pragma unroll
    for (int q = 1; q < kQ; ++q) St<T>::st(g + off[q], v[q]);
  }
}

template <typename T, bool ODD>
__global__ void __launch_bounds__(kBlock) k_macros(const lbm3d_params p, const T *__restrict__ g,
                                                   const int32_t *__restrict__ nbr,
                                                   const int32_t *__restrict__ face_tag,
                                                   const float *__restrict__ face_u, int first,
                                                   float *__restrict__ rho_out,
                                                   float *__restrict__ v_out) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)p.n_cells;
  if (x >= N) return;
  float v[kQ];
  uint32_t off[kQ];
  read_cell<T, ODD>(g, nbr, x, N, v, off);
  if (x >= (uint32_t)p.n_interior && !first) {
    const uint32_t fi = x - (uint32_t)p.n_interior;
    face_populations(v, face_tag[fi], face_u, fi, N - (uint32_t)p.n_interior, p);
  }
  float rho, ux, uy, uz;
  macros_of(v, p, first != 0, rho, ux, uy, uz);
  rho_out[x] = rho;
  v_out[x] = ux;
  v_out[(size_t)N + x] = uy;
  v_out[2 * (size_t)N + x] = uz;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_init(const lbm3d_params p, T *__restrict__ g,
                                                 float *__restrict__ face_u,
                                                 const float *__restrict__ rho0,
                                                 const float *__restrict__ u0) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)p.n_cells;
  if (x >= N) return;
  float rho = 1.0f, ux = 0.f, uy = 0.f, uz = 0.f;
  if (rho0) rho = rho0[x];
  if (u0) {
    ux = u0[x];
    uy = u0[(size_t)N + x];
    uz = u0[2 * (size_t)N + x];
  }
  const float uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int q = 0; q < kQ; ++q) St<T>::st(g + (size_t)q * N + x, feq_shift(q, rho, ux, uy, uz, uu));
  if (x >= (uint32_t)p.n_interior) {
    const uint32_t fi = x - (uint32_t)p.n_interior, nf = N - (uint32_t)p.n_interior;
    face_u[fi] = ux;
    face_u[nf + fi] = uy;
    face_u[2 * nf + fi] = uz;
  }
}

// ---------------------------------------------------------------------------
// Compaction.  Tiles of kTile cells (kItems per thread); per tile the counts
// of interior and face-owned fluid cells, an exclusive scan over tiles, then a
// block-level scan inside each tile to number the cells.
// ---------------------------------------------------------------------------
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;

// class of dense cell c: 0 solid, 1 interior fluid, 2 face-owned fluid (tag set)
__device__ __forceinline__ int classify(const lbm3d_params &p, const int8_t *__restrict__ solid,
                                        int64_t c, int32_t &tag) {
  if (solid[c] != 0) return 0;  // fluid iff solid == 0 (SP:225)
  const int nyz = p.ny * p.nz;
  const int i = (int)(c / nyz);
  const int j = (int)((c / p.nz) % p.ny);
  const int k = (int)(c % p.nz);
  int owner = -1;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int m = p.bc_mode[f];
    if (m != 1 && m != 2) continue;
    const int a = f >> 1;
    const int pos = a == 0 ? i : (a == 1 ? j : k);
    const int n = a == 0 ? p.nx : (a == 1 ? p.ny : p.nz);
    if (pos == ((f & 1) ? n - 1 : 0)) owner = f;  // later faces overwrite earlier ones
  }
  if (owner < 0) return 1;
  const int a = owner >> 1;
  int ii = i, jj = j, kk = k;
  const int n = a == 0 ? p.nx : (a == 1 ? p.ny : p.nz);
  const int inw = (owner & 1) ? n - 2 : 1;
  if (a == 0) ii = inw; else if (a == 1) jj = inw; else kk = inw;
  const int inward_solid = solid[((int64_t)ii * p.ny + jj) * p.nz + kk] > 0 ? 1 : 0;  // SP:278
  tag = owner | (p.bc_mode[owner] << 3) | (inward_solid << 5);
  return 2;
}

__global__ void __launch_bounds__(kBlock) k_compact_count(const lbm3d_params p,
                                                          const int8_t *__restrict__ solid,
                                                          int64_t n_dense, int32_t *__restrict__ tile_cnt) {
  __shared__ int s_cnt[2];
  if (threadIdx.x == 0) s_cnt[0] = s_cnt[1] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int ci = 0, cf = 0;
  for (int it = 0; it < kItems; ++it) {
    const int64_t c = base + (int64_t)threadIdx.x * kItems + it;
    if (c >= n_dense) break;
    int32_t tag;
    const int cls = classify(p, solid, c, tag);
    ci += cls == 1;
    cf += cls == 2;
  }
  if (ci) atomicAdd(&s_cnt[0], ci);
  if (cf) atomicAdd(&s_cnt[1], cf);
  __syncthreads();
  if (threadIdx.x == 0) {
    tile_cnt[2 * blockIdx.x] = s_cnt[0];
    tile_cnt[2 * blockIdx.x + 1] = s_cnt[1];
  }
}

// inclusive scan of one int per thread across a kBlock-thread block
__device__ __forceinline__ int block_inclusive_scan(int val, int *s_warp) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(val, d, 64);
    if (lane >= d) val += o;
  }
  if (lane == 63) s_warp[wid] = val;
  __syncthreads();
  int add = 0;
  for (int w = 0; w < wid; ++w) add += s_warp[w];
  __syncthreads();
  return val + add;
}

// exclusive scan over tiles (single block, sequential chunks), totals to counts
__global__ void __launch_bounds__(kBlock) k_compact_scan(int32_t *__restrict__ tile_cnt, int64_t n_tiles,
                                                         int32_t *__restrict__ counts) {
  __shared__ int s_warp[kBlock / 64];
  int carry_i = 0, carry_f = 0;
  for (int64_t start = 0; start < n_tiles; start += kBlock) {
    const int64_t t = start + threadIdx.x;
    const int vi = t < n_tiles ? tile_cnt[2 * t] : 0;
    const int vf = t < n_tiles ? tile_cnt[2 * t + 1] : 0;
    const int si = block_inclusive_scan(vi, s_warp);
    const int sf = block_inclusive_scan(vf, s_warp);
    if (t < n_tiles) {
      tile_cnt[2 * t] = carry_i + si - vi;
      tile_cnt[2 * t + 1] = carry_f + sf - vf;
    }
    __shared__ int s_tot[2];
    if (threadIdx.x == kBlock - 1) {
      s_tot[0] = si;
      s_tot[1] = sf;
    }
    __syncthreads();
    carry_i += s_tot[0];
    carry_f += s_tot[1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[0] = carry_i;
    counts[1] = carry_f;
  }
}

__global__ void __launch_bounds__(kBlock) k_compact_fill(const lbm3d_params p,
                                                         const int8_t *__restrict__ solid,
                                                         int64_t n_dense,
                                                         const int32_t *__restrict__ tile_off,
                                                         int32_t *__restrict__ d2c,
                                                         int32_t *__restrict__ cells,
                                                         int32_t *__restrict__ face_tag) {
  __shared__ int s_warp[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int cls[kItems];
  int32_t tags[kItems];
  int ci = 0, cf = 0;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int64_t c = base + (int64_t)threadIdx.x * kItems + it;
    cls[it] = 0;
    tags[it] = 0;
    if (c < n_dense) cls[it] = classify(p, solid, c, tags[it]);
    ci += cls[it] == 1;
    cf += cls[it] == 2;
  }
  int oi = block_inclusive_scan(ci, s_warp) - ci + tile_off[2 * blockIdx.x];
  int of = block_inclusive_scan(cf, s_warp) - cf + tile_off[2 * blockIdx.x + 1];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int64_t c = base + (int64_t)threadIdx.x * kItems + it;
    if (c >= n_dense) break;
    int32_t id = -1;
    if (cls[it] == 1) {
      id = oi++;
    } else if (cls[it] == 2) {
      face_tag[of] = tags[it];
      id = p.n_interior + of++;
    }
    if (d2c) d2c[c] = id;
    if (id >= 0) cells[id] = (int32_t)c;
  }
}

__global__ void __launch_bounds__(kBlock) k_neighbors(const lbm3d_params p, const int32_t *__restrict__ d2c,
                                                      const int32_t *__restrict__ cells,
                                                      int32_t *__restrict__ nbr) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)p.n_cells;
  if (x >= N) return;
  const int64_t c = cells[x];
  const int nyz = p.ny * p.nz;
  const int i = (int)(c / nyz), j = (int)((c / p.nz) % p.ny), k = (int)(c % p.nz);
#pragma unroll
  for (int q = 1; q < kQ; ++q) {
    // periodic_index, SP:247-257: -1 -> n-1, n -> 0 on every axis
    int a = i + ex(q), b = j + ey(q), d = k + ez(q);
    a = a < 0 ? p.nx - 1 : (a > p.nx - 1 ? 0 : a);
    b = b < 0 ? p.ny - 1 : (b > p.ny - 1 ? 0 : b);
    d = d < 0 ? p.nz - 1 : (d > p.nz - 1 ? 0 : d);
    nbr[(size_t)(q - 1) * N + x] = d2c[((int64_t)a * p.ny + b) * p.nz + d];
  }
}

__global__ void __launch_bounds__(kBlock) k_scatter(const lbm3d_params p, const int32_t *__restrict__ cells,
                                                    const float *__restrict__ rho,
                                                    const float *__restrict__ v, int32_t ox, int32_t oy,
                                                    int32_t oz, float *__restrict__ rho_d,
                                                    float *__restrict__ v_d) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)p.n_cells;
  if (x >= N) return;
  const int64_t c = cells[x];
  const int nyz = p.ny * p.nz;
  const int i = (int)(c / nyz), j = (int)((c / p.nz) % p.ny), k = (int)(c % p.nz);
  const int64_t o = ((int64_t)i * oy + j) * oz + k;
  if (rho_d) rho_d[o] = rho[x];
  if (v_d) {
    v_d[3 * o] = v[x];
    v_d[3 * o + 1] = v[(size_t)N + x];
    v_d[3 * o + 2] = v[2 * (size_t)N + x];
  }
}

// FP64 reductions: per-block partials, then one block folds them.
constexpr int kRed = 8;
__global__ void __launch_bounds__(kBlock) k_reduce_partial(const lbm3d_params p,
                                                           const int32_t *__restrict__ cells,
                                                           const float *__restrict__ rho,
                                                           const float *__restrict__ v, int32_t i_lo,
                                                           int32_t i_hi, double *__restrict__ part) {
  const uint32_t N = (uint32_t)p.n_cells;
  double acc[kRed] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nyz = p.ny * p.nz;
  for (uint32_t x = blockIdx.x * kBlock + threadIdx.x; x < N; x += gridDim.x * kBlock) {
    const float r = rho[x], a = v[x], b = v[(size_t)N + x], c = v[2 * (size_t)N + x];
    const bool fin = isfinite(r) && isfinite(a) && isfinite(b) && isfinite(c);
    if (!fin) {
      acc[5] += 1.0;
      continue;
    }
    acc[0] += r;
    acc[1] += a;
    acc[2] += b;
    acc[3] += c;
    const double nrm = sqrt((double)a * a + (double)b * b + (double)c * c);
    acc[4] = nrm > acc[4] ? nrm : acc[4];
    const int i = (int)(cells[x] / nyz);
    if (i >= i_lo && i <= i_hi) {
      acc[6] += a;
      acc[7] += 1.0;
    }
  }
  __shared__ double s[kRed][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < kRed; ++r) {
    double val = acc[r];
    for (int d = 32; d > 0; d >>= 1) {
      const double o = __shfl_down(val, d, 64);
      val = r == 4 ? (o > val ? o : val) : val + o;
    }
    if (lane == 0) s[r][wid] = val;
  }
  __syncthreads();
  if (threadIdx.x < kRed) {
    const int r = threadIdx.x;
    double val = s[r][0];
    for (int w = 1; w < kBlock / 64; ++w) val = r == 4 ? (s[r][w] > val ? s[r][w] : val) : val + s[r][w];
    part[(size_t)blockIdx.x * kRed + r] = val;
  }
}

__global__ void __launch_bounds__(kBlock) k_reduce_final(const double *__restrict__ part, int nparts,
                                                         double *__restrict__ out) {
  if (threadIdx.x >= kRed) return;
  const int r = threadIdx.x;
  double val = 0.0;
  for (int b = 0; b < nparts; ++b) {
    const double o = part[(size_t)b * kRed + r];
    val = r == 4 ? (o > val ? o : val) : val + o;
  }
  out[r] = val;
}

constexpr int kRedBlocks = 1024;

int check_params(const lbm3d_params *p) {
  if (!p) return fail(LBM3D_EINVAL, "%s", "params is NULL");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(LBM3D_EINVAL, "%s", "grid extents must be >= 1");
  if ((int64_t)p->nx * p->ny * p->nz > INT32_MAX)
    return fail(LBM3D_ESIZE, "%s", "grid has more than 2^31-1 cells");
  for (int f = 0; f < 6; ++f) {
    if (p->bc_mode[f] == 1 || p->bc_mode[f] == 2) {
      const int n = f < 2 ? p->nx : (f < 4 ? p->ny : p->nz);
      if (n < 2) return fail(LBM3D_EINVAL, "%s", "a face with a boundary condition needs extent >= 2");
    }
  }
  return LBM3D_OK;
}

int check_cells(const lbm3d_params *p) {
  if (p->n_cells < 0 || p->n_interior < 0 || p->n_interior > p->n_cells)
    return fail(LBM3D_EINVAL, "%s", "n_cells / n_interior out of range");
  if ((uint64_t)p->n_cells * kQ > UINT32_MAX)
    return fail(LBM3D_ESIZE, "%s", "n_cells * 19 exceeds the 32-bit slot offsets");
  return LBM3D_OK;
}

template <typename T>
int launch_step(hipStream_t s, const lbm3d_params &p, void *g, const int32_t *nbr, const int32_t *tag,
                float *face_u, int64_t t) {
  const unsigned grid = (unsigned)((p.n_cells + kBlock - 1) / kBlock);
  if (grid == 0) return LBM3D_OK;
  if (t & 1)
    k_step<T, true><<<grid, kBlock, 0, s>>>(p, (T *)g, nbr, tag, face_u, t == 0);
  else
    k_step<T, false><<<grid, kBlock, 0, s>>>(p, (T *)g, nbr, tag, face_u, t == 0);
  return check_launch("lbm3d_step");
}

template <typename T>
int launch_macros(hipStream_t s, const lbm3d_params &p, const void *g, const int32_t *nbr,
                  const int32_t *tag, const float *face_u, int64_t t, float *rho, float *v) {
  const unsigned grid = (unsigned)((p.n_cells + kBlock - 1) / kBlock);
  if (grid == 0) return LBM3D_OK;
  if (t & 1)
    k_macros<T, true><<<grid, kBlock, 0, s>>>(p, (const T *)g, nbr, tag, face_u, t == 0, rho, v);
  else
    k_macros<T, false><<<grid, kBlock, 0, s>>>(p, (const T *)g, nbr, tag, face_u, t == 0, rho, v);
  return check_launch("lbm3d_macros");
}

template <typename T>
int launch_init(hipStream_t s, const lbm3d_params &p, void *g, float *face_u, const float *rho0,
                const float *u0) {
  const unsigned grid = (unsigned)((p.n_cells + kBlock - 1) / kBlock);
  if (grid == 0) return LBM3D_OK;
  k_init<T><<<grid, kBlock, 0, s>>>(p, (T *)g, face_u, rho0, u0);
  return check_launch("lbm3d_init");
}

bool valid_dtype(int d) { return d == LBM3D_F32 || d == LBM3D_F16 || d == LBM3D_BF16; }

int solver_args(const lbm3d_params *p, int dtype, const void *g, const int32_t *nbr,
                const int32_t *tag, const float *face_u) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!valid_dtype(dtype)) return fail(LBM3D_EINVAL, "%s", "unknown dtype");
  if (p->n_cells > 0 && (!g || !nbr)) return fail(LBM3D_EINVAL, "%s", "populations or neighbour table is NULL");
  if (p->n_cells > p->n_interior && (!tag || !face_u))
    return fail(LBM3D_EINVAL, "%s", "face cells present but face_tag or face_u is NULL");
  return LBM3D_OK;
}

}  // namespace

extern "C" {

int lbm3d_abi_version(void) { return LBM3D_ABI_VERSION; }

const char *lbm3d_last_error(void) { return g_err; }

size_t lbm3d_compact_workspace_bytes(int64_t n_dense) {
  const int64_t tiles = (n_dense + kTile - 1) / kTile;
  return (size_t)(tiles > 0 ? tiles : 1) * 2 * sizeof(int32_t);
}

int lbm3d_compact_count(void *stream, const lbm3d_params *p, const int8_t *solid, int32_t *counts,
                        void *workspace, size_t workspace_bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!solid || !counts || !workspace) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_compact_count");
  const int64_t n = (int64_t)p->nx * p->ny * p->nz;
  if (workspace_bytes < lbm3d_compact_workspace_bytes(n))
    return fail(LBM3D_EINVAL, "%s", "compaction workspace too small");
  const int64_t tiles = (n + kTile - 1) / kTile;
  hipStream_t s = (hipStream_t)stream;
  k_compact_count<<<(unsigned)tiles, kBlock, 0, s>>>(*p, solid, n, (int32_t *)workspace);
  if ((rc = check_launch("lbm3d_compact_count"))) return rc;
  k_compact_scan<<<1, kBlock, 0, s>>>((int32_t *)workspace, tiles, counts);
  return check_launch("lbm3d_compact_count/scan");
}

int lbm3d_compact_fill(void *stream, const lbm3d_params *p, const int8_t *solid, int32_t *d2c,
                       int32_t *cells, int32_t *face_tag, void *workspace, size_t workspace_bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!solid || !workspace || (p->n_cells > 0 && !cells) || (p->n_cells > p->n_interior && !face_tag))
    return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_compact_fill");
  const int64_t n = (int64_t)p->nx * p->ny * p->nz;
  if (workspace_bytes < lbm3d_compact_workspace_bytes(n))
    return fail(LBM3D_EINVAL, "%s", "compaction workspace too small");
  const int64_t tiles = (n + kTile - 1) / kTile;
  k_compact_fill<<<(unsigned)tiles, kBlock, 0, (hipStream_t)stream>>>(*p, solid, n, (const int32_t *)workspace,
                                                                      d2c, cells, face_tag);
  return check_launch("lbm3d_compact_fill");
}

int lbm3d_build_neighbors(void *stream, const lbm3d_params *p, const int32_t *d2c, const int32_t *cells,
                          int32_t *nbr) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (p->n_cells == 0) return LBM3D_OK;
  if (!d2c || !cells || !nbr) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_build_neighbors");
  const unsigned grid = (unsigned)((p->n_cells + kBlock - 1) / kBlock);
  k_neighbors<<<grid, kBlock, 0, (hipStream_t)stream>>>(*p, d2c, cells, nbr);
  return check_launch("lbm3d_build_neighbors");
}

int lbm3d_init(void *stream, const lbm3d_params *p, int dtype, void *g, float *face_u, const float *rho0,
               const float *u0) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!valid_dtype(dtype)) return fail(LBM3D_EINVAL, "%s", "unknown dtype");
  if (p->n_cells > 0 && !g) return fail(LBM3D_EINVAL, "%s", "populations are NULL");
  if (p->n_cells > p->n_interior && !face_u) return fail(LBM3D_EINVAL, "%s", "face_u is NULL");
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case LBM3D_F32: return launch_init<float>(s, *p, g, face_u, rho0, u0);
    case LBM3D_F16: return launch_init<__half>(s, *p, g, face_u, rho0, u0);
    default: return launch_init<__hip_bfloat16>(s, *p, g, face_u, rho0, u0);
  }
}

int lbm3d_step(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *nbr,
               const int32_t *face_tag, float *face_u, int64_t t) {
  int rc = solver_args(p, dtype, g, nbr, face_tag, face_u);
  if (rc) return rc;
  if (t < 0) return fail(LBM3D_EINVAL, "%s", "negative step index");
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case LBM3D_F32: return launch_step<float>(s, *p, g, nbr, face_tag, face_u, t);
    case LBM3D_F16: return launch_step<__half>(s, *p, g, nbr, face_tag, face_u, t);
    default: return launch_step<__hip_bfloat16>(s, *p, g, nbr, face_tag, face_u, t);
  }
}

int lbm3d_run(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *nbr,
              const int32_t *face_tag, float *face_u, int64_t t0, int32_t nsteps) {
  int rc = solver_args(p, dtype, g, nbr, face_tag, face_u);
  if (rc) return rc;
  if (t0 < 0 || nsteps < 0) return fail(LBM3D_EINVAL, "%s", "negative step index or count");
  hipStream_t s = (hipStream_t)stream;
  for (int32_t k = 0; k < nsteps; ++k) {
    switch (dtype) {
      case LBM3D_F32: rc = launch_step<float>(s, *p, g, nbr, face_tag, face_u, t0 + k); break;
      case LBM3D_F16: rc = launch_step<__half>(s, *p, g, nbr, face_tag, face_u, t0 + k); break;
      default: rc = launch_step<__hip_bfloat16>(s, *p, g, nbr, face_tag, face_u, t0 + k); break;
    }
    if (rc) return rc;
  }
  return LBM3D_OK;
}

int lbm3d_macros(void *stream, const lbm3d_params *p, int dtype, const void *g, const int32_t *nbr,
                 const int32_t *face_tag, const float *face_u, int64_t t, float *rho, float *v) {
  int rc = solver_args(p, dtype, g, nbr, face_tag, face_u);
  if (rc) return rc;
  if (t < 0) return fail(LBM3D_EINVAL, "%s", "negative step index");
  if (p->n_cells > 0 && (!rho || !v)) return fail(LBM3D_EINVAL, "%s", "rho or v output is NULL");
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case LBM3D_F32: return launch_macros<float>(s, *p, g, nbr, face_tag, face_u, t, rho, v);
    case LBM3D_F16: return launch_macros<__half>(s, *p, g, nbr, face_tag, face_u, t, rho, v);
    default: return launch_macros<__hip_bfloat16>(s, *p, g, nbr, face_tag, face_u, t, rho, v);
  }
}

int lbm3d_scatter_dense(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho,
                        const float *v, int32_t ox, int32_t oy, int32_t oz, float *rho_dense,
                        float *v_dense) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (ox < p->nx || oy < p->ny || oz < p->nz) return fail(LBM3D_EINVAL, "%s", "dense extent smaller than the grid");
  if (p->n_cells == 0) return LBM3D_OK;
  if (!cells || (rho_dense && !rho) || (v_dense && !v)) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_scatter_dense");
  const unsigned grid = (unsigned)((p->n_cells + kBlock - 1) / kBlock);
  k_scatter<<<grid, kBlock, 0, (hipStream_t)stream>>>(*p, cells, rho, v, ox, oy, oz, rho_dense, v_dense);
  return check_launch("lbm3d_scatter_dense");
}

size_t lbm3d_reduce_workspace_bytes(int32_t n_cells) {
  (void)n_cells;
  return (size_t)kRedBlocks * kRed * sizeof(double);
}

int lbm3d_reduce(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho, const float *v,
                 int32_t i_lo, int32_t i_hi, double *out, void *workspace, size_t workspace_bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!out || !workspace || (p->n_cells > 0 && (!cells || !rho || !v)))
    return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_reduce");
  if (workspace_bytes < lbm3d_reduce_workspace_bytes(p->n_cells))
    return fail(LBM3D_EINVAL, "%s", "reduction workspace too small");
  int blocks = (p->n_cells + kBlock - 1) / kBlock;
  blocks = blocks < 1 ? 1 : (blocks > kRedBlocks ? kRedBlocks : blocks);
  hipStream_t s = (hipStream_t)stream;
  k_reduce_partial<<<blocks, kBlock, 0, s>>>(*p, cells, rho, v, i_lo, i_hi, (double *)workspace);
  if ((rc = check_launch("lbm3d_reduce"))) return rc;
  k_reduce_final<<<1, 64, 0, s>>>((const double *)workspace, blocks, out);
  return check_launch("lbm3d_reduce/final");
}

}  // extern "C"
