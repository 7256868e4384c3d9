// liblbm3d.so -- single-phase sparse D3Q19 MRT lattice Boltzmann for MI355X (gfx950).
//
// Replaces the Taichi kernels of the reference class LB3D_Solver_Single_Phase
// (Single_phase/LBM_3D_SinglePhase_Solver.py, "SP"): its four-kernel step
// colission -> streaming1 -> Boundary_condition -> streaming3 (SP:477-481)
// becomes one fused kernel over a compacted fluid-cell list, in place with the
// AA access pattern, with the MRT transform held in registers.  The C ABI is
// declared in include/lbm3d.h.
//
// Data in HBM (n fluid cells in list order, S = n rounded up to 64):
//   g[19][S]       shifted populations f - w (FP32, FP16 or BF16)
//   cells[n]       C-order dense index of each fluid cell
//   rows[nx*ny][W] per grid row (i,j) and 32-cell word: fluid bit mask and the
//                  compact index of the word's first fluid cell (8 bytes)
//   face_bits      2 bits per cell: face-owned, owner's inward neighbour solid
//   face_u         previous-step velocity of pressure-face nodes, per face plane
// Neighbours are found through `rows` (compact index = base + popcount of the
// mask below the cell), so no per-cell neighbour table is stored or read.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <type_traits>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "lbm3d.h"
#include "xcd.h"

namespace {

constexpr int kQ = 19;
#ifndef LBM3D_BLOCK
#define LBM3D_BLOCK 256
#endif
constexpr int kBlock = LBM3D_BLOCK;  // threads per block, a multiple of 64
constexpr int kWinWords = 384;       // LDS row window: 3 planes x (J+3) rows x W words

// Build-time knobs (defaults are the measured best on MI355X).
#ifndef LBM3D_XCD_CHUNK_EVEN
#define LBM3D_XCD_CHUNK_EVEN 128  // C > 0: each XCD takes runs of C consecutive blocks on even steps
#endif
#ifndef LBM3D_XCD_CHUNK_ODD
// ... and on odd steps (0: the hardware's round robin).  An odd step's cells
// read and write slot lines shared with the cells one plane away (R rows of
// cells away in a list in bands of R rows) and the bounce-back slots of their
// own lines; runs of consecutive blocks keep most of those sharers on one
// XCD's L2.  On a 256^3 pack in bands of 32 rows, runs of 128 blocks on both
// parities: 166 us per step against 169-170 with runs of 256 on odd steps only.
#define LBM3D_XCD_CHUNK_ODD 128
#endif
// The batch kernels' runs: 256 blocks on odd steps, the hardware's round
// robin on even steps.  A batch group's slots stay in the last-level cache
// across its steps, and runs of 128 or 256 blocks on both parities measured
// the same to 0.5 % on 64 packs.
#ifndef LBM3D_BATCH_XCD_CHUNK_EVEN
#define LBM3D_BATCH_XCD_CHUNK_EVEN 0
#endif
#ifndef LBM3D_BATCH_XCD_CHUNK_ODD
#define LBM3D_BATCH_XCD_CHUNK_ODD 256
#endif
#ifndef LBM3D_STEP_MINB
#define LBM3D_STEP_MINB 1
#endif
#ifndef LBM3D_STEP_WAVES16
// Minimum waves per SIMD of the FP16/BF16 step kernels.  Left alone, their odd
// step takes 74 VGPRs (6 waves); held to 7 waves it fits 72 without spilling:
// FP16 85.3 / 86.4 us per step, FP16 with stochastic rounding 98.6 / 101.4,
// BF16 with stochastic rounding 92.1 / 92.9 (8 waves spill 16-28 bytes and lose).
#define LBM3D_STEP_WAVES16 7
#endif
#ifndef LBM3D_REVERSE_EVEN
// Even steps walk the block range backwards, so they first read the slots the
// preceding odd step wrote last, while those are still in the 256 MB last-level
// cache; the next odd step, walking forwards, likewise starts where this one ended.
#define LBM3D_REVERSE_EVEN 1
#endif
#ifndef LBM3D_ROW_WINDOW
#define LBM3D_ROW_WINDOW 1    // odd steps read row words from a per-workgroup LDS window
#endif
#ifndef LBM3D_CPOL_LOAD
#define LBM3D_CPOL_LOAD 0     // cache-policy bits of population loads
#endif
#ifndef LBM3D_CPOL_STORE
#define LBM3D_CPOL_STORE 0    // cache-policy bits of population stores
#endif

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
template <typename R = float>
__host__ __device__ constexpr R wq(int q) {
  return q == 0 ? R(1) / R(3) : (q < 7 ? R(1) / R(18) : R(1) / R(36));
}

// slot stride: n rounded up to whole 256-byte lines of FP32
__host__ __device__ constexpr uint32_t slot_stride(uint32_t n) { return (n + 63u) & ~63u; }

// ---------------------------------------------------------------------------
// Kernel parameters: the public struct plus values derived once on the host.
// ---------------------------------------------------------------------------
// Real parameters in the kernel's arithmetic type, rounded once on the host.
template <typename R>
struct Phys {
  R sv, so, fx, fy, fz, ga, gb;
  R rho_bc[6];
  R u_bc[6][3];
};

// Grid geometry of one sample: everything the kernels derive from (nx, ny, nz).
struct Geo {
  int32_t nx, ny, nz;
  uint32_t W;              // 32-cell words per grid row
  uint32_t nyz;            // ny * nz
  uint32_t m_nz, s_nz;     // c / nz  == (c * m_nz) >> s_nz   for c < 2^31
  uint32_t m_nyz, s_nyz;   // c / nyz == (c * m_nyz) >> s_nyz
  uint32_t plane[6];       // cells per face plane
  int32_t face_u_off[6];   // first face_u element of each pressure face's plane
  uint32_t n_row_words;    // nx * ny * W
  uint32_t band;           // rows j per band of the fluid list's order, in [1, ny] (ny: C order)
};

struct KP {
  lbm3d_params p;          // integer fields are read on the device; its doubles are not
  Phys<float> pf;
  Phys<double> pd;
  Geo geo;
  uint32_t S;              // slot stride
  uint32_t sr_seed;        // per-step seed of stochastic rounding, 0 = round to nearest
};

// Exact division of 31-bit numerators by d: m = ceil(2^(31+l) / d), l = ceil(log2 d).
void magic(uint32_t d, uint32_t &m, uint32_t &s) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t two = 1ull << (31 + l);
  m = (uint32_t)((two + d - 1) / d);
  s = 31 + l;
}

Geo make_geo(int32_t nx, int32_t ny, int32_t nz, const int32_t face_u_off[6], int32_t band_rows) {
  Geo g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.W = (uint32_t)(nz + 31) / 32;
  g.nyz = (uint32_t)ny * (uint32_t)nz;
  magic((uint32_t)nz, g.m_nz, g.s_nz);
  magic(g.nyz, g.m_nyz, g.s_nyz);
  g.plane[0] = g.plane[1] = (uint32_t)ny * nz;
  g.plane[2] = g.plane[3] = (uint32_t)nx * nz;
  g.plane[4] = g.plane[5] = (uint32_t)nx * ny;
  for (int f = 0; f < 6; ++f) g.face_u_off[f] = face_u_off[f];
  g.n_row_words = (uint32_t)nx * (uint32_t)ny * g.W;
  g.band = (uint32_t)(band_rows <= 0 || band_rows >= ny ? ny : band_rows);
  return g;
}

template <typename R>
void fill_phys(Phys<R> &f, const lbm3d_params &p) {
  f.sv = (R)p.s_v;
  f.so = (R)p.s_o;
  f.fx = (R)p.fx;
  f.fy = (R)p.fy;
  f.fz = (R)p.fz;
  f.ga = (R)p.guo_a;
  f.gb = (R)p.guo_b;
  for (int i = 0; i < 6; ++i) {
    f.rho_bc[i] = (R)p.rho_bc[i];
    for (int c = 0; c < 3; ++c) f.u_bc[i][c] = (R)p.u_bc[i][c];
  }
}

KP make_kp(const lbm3d_params &p) {
  KP k;
  k.p = p;
  fill_phys(k.pf, p);
  fill_phys(k.pd, p);
  k.S = slot_stride((uint32_t)(p.n_cells > 0 ? p.n_cells : 0));
  k.geo = make_geo(p.nx, p.ny, p.nz, p.face_u_off, p.band_rows);
  k.sr_seed = 0u;
  return k;
}

template <typename R> __device__ __forceinline__ const Phys<R> &phys(const KP &kp);
template <> __device__ __forceinline__ const Phys<float> &phys<float>(const KP &kp) { return kp.pf; }
template <> __device__ __forceinline__ const Phys<double> &phys<double>(const KP &kp) { return kp.pd; }

__device__ __forceinline__ uint32_t fdiv(uint32_t c, uint32_t m, uint32_t s) {
  return (uint32_t)(((uint64_t)c * m) >> s);
}

__device__ __forceinline__ void coords(const Geo &gm, uint32_t c, int &i, int &j, int &k) {
  const uint32_t a = fdiv(c, gm.m_nyz, gm.s_nyz);
  const uint32_t r = c - a * gm.nyz;
  const uint32_t b = fdiv(r, gm.m_nz, gm.s_nz);
  i = (int)a;
  j = (int)b;
  k = (int)(r - b * (uint32_t)gm.nz);
}

// periodic_index, SP:247-257: -1 -> n-1 and n -> 0 on every axis, whatever the BC mode
__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? n - 1 : (a > n - 1 ? 0 : a); }

// Block -> cell-range map: plain, or runs of C blocks per XCD (xcd.h).
// this workgroup's block among nb (nb: none); hardware block h is the
// (h / 8)-th block of XCD h % 8
template <uint32_t C, bool REVERSE = false>
__device__ __forceinline__ uint32_t logical_block(uint32_t nb) {
  const uint32_t h = blockIdx.x;
  uint32_t b = h;
  if constexpr (C > 0) b = xcd_block(C, h & 7u, h >> 3, nb);
  if (b >= nb) return nb;
  return REVERSE ? nb - 1u - b : b;
}

// ---------------------------------------------------------------------------
// Row lookup.  rows[(i*ny + j)*W + w] covers cells k = 32w .. 32w+31 of grid
// row (i, j): bit b of .x set when cell 32w+b is fluid, .y the compact id of
// the word's first fluid cell, .z / .w the compact ids (or -1 when not fluid)
// of the cells just before the word's first cell and just after its last one,
// with the periodic wrap of periodic_index (SP:247-257) at k = 0 and nz-1.
// Every row's fluid cells have consecutive ids in k order (in either list
// order), so inside a word the id of cell k is .y plus the number of fluid
// cells below it.
// ---------------------------------------------------------------------------
struct Rows {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Rows(const uint4 *rows, uint32_t n_words) {
    r = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4 *>(rows), (short)0, (int)(n_words * 16u), 0x00020000);
  }
  __device__ __forceinline__ uint4 word(uint32_t row, uint32_t W, int k) const {
    const uint32_t off = (row * W + ((uint32_t)k >> 5)) * 16u;
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
  }
};

__device__ __forceinline__ int32_t id_in_word(uint4 w, int k) {
  const uint32_t b = (uint32_t)k & 31u;
  const bool fl = (w.x >> b) & 1u;
  return fl ? (int32_t)(w.y + __popc(w.x & ((1u << b) - 1u))) : -1;
}

// ids of cells (k-1, k, k+1) of one row, k-1 and k+1 wrapped, from its word w
__device__ __forceinline__ void row3(const uint4 w, const Geo &gm, int k, int32_t &lo, int32_t &mid, int32_t &hi) {
  const uint32_t b = (uint32_t)k & 31u;
  const uint32_t below = w.y + __popc(w.x & ((1u << b) - 1u));
  const uint32_t fm = (w.x >> b) & 1u;
  mid = fm ? (int32_t)below : -1;
  const bool last = b == 31u || k == gm.nz - 1;
  const bool up_fluid = (w.x >> ((b + 1u) & 31u)) & 1u;
  hi = last ? (int32_t)w.w : (up_fluid ? (int32_t)(below + fm) : -1);
  const bool down_fluid = (w.x >> ((b - 1u) & 31u)) & 1u;
  lo = b == 0u ? (int32_t)w.z : (down_fluid ? (int32_t)below - 1 : -1);
}

// The 9 row words around a cell (i, j): word(di, dj, k) is the word holding k
// of row (i + di, j + dj), wrapped (periodic_index, SP:247-257).
// From the row table in memory:
struct TableRows {
  const Rows &rb;
  uint32_t W, ny, ri[3], rj[3];
  __device__ __forceinline__ TableRows(const Rows &r, const Geo &gm, int i, int j) : rb(r) {
    W = gm.W;
    ny = (uint32_t)gm.ny;
    ri[0] = (uint32_t)wrap(i - 1, gm.nx);
    ri[1] = (uint32_t)i;
    ri[2] = (uint32_t)wrap(i + 1, gm.nx);
    rj[0] = (uint32_t)wrap(j - 1, gm.ny);
    rj[1] = (uint32_t)j;
    rj[2] = (uint32_t)wrap(j + 1, gm.ny);
  }
  __device__ __forceinline__ uint4 word(int di, int dj, int k) const {
    return rb.word(ri[di + 1] * ny + rj[dj + 1], W, k);
  }
};
// From a workgroup's LDS copy of the rows i0-1 .. i0+1 x j0-1 .. j0+J+1 (JW = J + 3
// rows per plane, W words per row), for cell k of row (i0, j0 + jrel): c points
// at the cell's own word, and the other 8 sit a wave-uniform stride away
// (si = JW W words per plane, sj = W per row).
struct WindowRows {
  const uint4 *c;
  int si, sj;
  __device__ __forceinline__ WindowRows(const uint4 *win, uint32_t W, uint32_t JW, int jrel, int k)
      : c(win + (((JW + (uint32_t)jrel + 1u) * W) + ((uint32_t)k >> 5))), si((int)(JW * W)), sj((int)W) {}
  __device__ __forceinline__ uint4 word(int di, int dj, int) const { return c[di * si + dj * sj]; }
};

// Compact ids of the 18 neighbours (index q-1), -1 where the neighbour is
// not fluid: the push target of streaming1 (SP:264-266) or its bounce-back
// (SP:267-268).
template <class Src>
__device__ __forceinline__ void neighbor_ids(const Src &src, const Geo &gm, int k, int32_t nb[18]) {
  int32_t a, b, c;
  // own row: q5 (0,0,1), q6 (0,0,-1)
  row3(src.word(0, 0, k), gm, k, a, b, c);
  nb[4] = c;
  nb[5] = a;
  // row (i+1, j): q1 (1,0,0), q11 (1,0,1), q13 (1,0,-1)
  row3(src.word(1, 0, k), gm, k, a, b, c);
  nb[0] = b;
  nb[10] = c;
  nb[12] = a;
  // row (i-1, j): q2 (-1,0,0), q12 (-1,0,-1), q14 (-1,0,1)
  row3(src.word(-1, 0, k), gm, k, a, b, c);
  nb[1] = b;
  nb[11] = a;
  nb[13] = c;
  // row (i, j+1): q3 (0,1,0), q15 (0,1,1), q17 (0,1,-1)
  row3(src.word(0, 1, k), gm, k, a, b, c);
  nb[2] = b;
  nb[14] = c;
  nb[16] = a;
  // row (i, j-1): q4 (0,-1,0), q16 (0,-1,-1), q18 (0,-1,1)
  row3(src.word(0, -1, k), gm, k, a, b, c);
  nb[3] = b;
  nb[15] = a;
  nb[17] = c;
  // diagonal rows: q7 (1,1,0), q8 (-1,-1,0), q9 (1,-1,0), q10 (-1,1,0)
  nb[6] = id_in_word(src.word(1, 1, k), k);
  nb[7] = id_in_word(src.word(-1, -1, k), k);
  nb[8] = id_in_word(src.word(1, -1, k), k);
  nb[9] = id_in_word(src.word(-1, 1, k), k);
}

// ---------------------------------------------------------------------------
// Population storage.  g = f - w in T; FP32 arithmetic for FP32/FP16/BF16
// storage, FP64 for FP64.  Every access goes through a buffer resource covering
// one opposite pair of slots (directions 2k-1 and 2k; slot 0 alone): a
// wave-uniform descriptor in SGPRs plus a 32-bit per-lane byte offset, so no
// per-access 64-bit address arithmetic.  The descriptor's record count bounds
// every access.
// ---------------------------------------------------------------------------
constexpr int kBufferFlags = 0x00020000;  // dword 3 of a raw buffer resource on gfx9 parts

template <typename T> struct St;
template <> struct St<float> {
  using R = float;
  static constexpr bool kStochastic = false;
  static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, LBM3D_CPOL_LOAD));
  }
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)vo, (int)so, LBM3D_CPOL_STORE);
  }
};
// Stochastic rounding to a 16-bit format with gfx950's conversions
// (v_cvt_sr_bf16_f32, v_cvt_sr_f16_f32): uniform random bits are added to the
// FP32 pattern below the target format's last place and the result is
// truncated, so the stored value equals the exact one in expectation and an
// increment far below the format's spacing is kept on average instead of
// always rounding away.  The instructions take the dither from the top of
// their 32-bit random word (BF16: its high 16 bits; FP16: its top 13 bits for
// normal values, more below), so the 16 random bits rnd (bits 0-15) go to bits
// 16-31.  Inf and NaN pass through (NaN as the quiet NaN).
typedef __bf16 lbm3d_bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 lbm3d_f16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned short bf16_sr(float v, uint32_t rnd) {
  lbm3d_bf16x2 z = {};
  z = __builtin_amdgcn_cvt_sr_bf16_f32(z, v, rnd << 16, false);
  return __builtin_bit_cast(unsigned short, z[0]);
}

__device__ __forceinline__ unsigned short half_sr(float v, uint32_t rnd) {
  lbm3d_f16x2 z = {};
  z = __builtin_amdgcn_cvt_sr_f16_f32(z, v, rnd << 16, false);
  return __builtin_bit_cast(unsigned short, z[0]);
}

template <> struct St<__half> {
  using R = float;
  static constexpr bool kStochastic = true;
  static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)vo, (int)so, LBM3D_CPOL_LOAD);
    return __half2float(__ushort_as_half(b));
  }
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v) {
    __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half_rn(v)), r, (int)vo, (int)so,
                                          LBM3D_CPOL_STORE);
  }
  static __device__ __forceinline__ void st_sr(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v,
                                               uint32_t rnd) {
    __builtin_amdgcn_raw_buffer_store_b16(half_sr(v, rnd), r, (int)vo, (int)so, LBM3D_CPOL_STORE);
  }
};
template <> struct St<__hip_bfloat16> {
  using R = float;
  static constexpr bool kStochastic = true;
  static __device__ __forceinline__ void st_sr(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v,
                                               uint32_t rnd) {
    __builtin_amdgcn_raw_buffer_store_b16(bf16_sr(v, rnd), r, (int)vo, (int)so, LBM3D_CPOL_STORE);
  }
  static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)vo, (int)so, LBM3D_CPOL_LOAD);
    return __uint_as_float((uint32_t)b << 16);
  }
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v) {
    const __hip_bfloat16 h = __float2bfloat16(v);  // round to nearest even
    __builtin_amdgcn_raw_buffer_store_b16(*reinterpret_cast<const unsigned short *>(&h), r, (int)vo, (int)so,
                                          LBM3D_CPOL_STORE);
  }
};
template <> struct St<double> {
  using R = double;
  static constexpr bool kStochastic = false;
  static __device__ __forceinline__ double ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, LBM3D_CPOL_LOAD);
    return __hiloint2double((int)v[1], (int)v[0]);
  }
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)) w;
    w[0] = (uint32_t)b;
    w[1] = (uint32_t)(b >> 32);
    __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)vo, (int)so, LBM3D_CPOL_STORE);
  }
};

// 32-bit mixing function (a "lowbias32" integer hash)
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

// xorshift32 step (Marsaglia): the per-store random bits of one cell's stream
__device__ __forceinline__ uint32_t xorshift32(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}

// Resource k covers slots 2k-1 and 2k (k = 1..9); resource 0 covers slot 0.
template <typename T>
struct Slots {
  using R = typename St<T>::R;
  __amdgpu_buffer_rsrc_t r[10];
  uint32_t Sb;    // slot stride in bytes
  uint32_t seed;  // nonzero: stochastic rounding of 16-bit stores (the step's seed)
  __device__ __forceinline__ Slots(const T *g, uint32_t S, uint32_t sr_seed = 0u) {
    seed = sr_seed;
    Sb = S * (uint32_t)sizeof(T);
    char *base = (char *)const_cast<T *>(g);
    r[0] = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)Sb, kBufferFlags);
#pragma unroll
    for (int k = 1; k <= 9; ++k)
      r[k] = __builtin_amdgcn_make_buffer_rsrc(base + (size_t)(2 * k - 1) * Sb, (short)0, (int)(2 * Sb),
                                               kBufferFlags);
  }
  // Start of the random stream of the cell with id `local` in its own grid
  // (a batch sample's ids count from its first cell): the bits of a store
  // depend only on the step, that id and the store's place in the cell's
  // fixed store order, not on where the cell sits in memory.
  __device__ __forceinline__ uint32_t stream(uint32_t local) const {
    if constexpr (St<T>::kStochastic) {
      return seed ? mix32(seed + local * 0x9E3779B9u) | 1u : 0u;  // nonzero: xorshift32 never leaves 0
    }
    return 0u;
  }
  // slot q at byte offset vo within the slot
  __device__ __forceinline__ R ld_slot(int q, uint32_t vo) const {
    return St<T>::ld(r[(q + 1) >> 1], vo, (q > 0 && !(q & 1)) ? Sb : 0u);
  }
  // rs: the cell's stream (advanced by a stochastically rounded store)
  // k: the store's place (0..18) among the cell's stores; each xorshift32 step
  // feeds two stores, 16 bits each
  __device__ __forceinline__ void store(int pair, uint32_t vo, uint32_t so, R v, uint32_t &rs, int k) const {
    if constexpr (St<T>::kStochastic) {
      if (seed) {
        if (!(k & 1)) rs = xorshift32(rs);
        St<T>::st_sr(r[pair], vo, so, v, (k & 1) ? rs >> 16 : rs & 0xFFFFu);
        return;
      }
    }
    St<T>::st(r[pair], vo, so, v);
  }
  __device__ __forceinline__ void st_slot(int q, uint32_t vo, R v, uint32_t &rs, int k) const {
    store((q + 1) >> 1, vo, (q > 0 && !(q & 1)) ? Sb : 0u, v, rs, k);
  }
  // byte offset vo relative to the pair of direction q (vo < 2 Sb)
  __device__ __forceinline__ R ld_pair(int q, uint32_t vo) const { return St<T>::ld(r[(q + 1) >> 1], vo, 0u); }
  __device__ __forceinline__ void st_pair(int q, uint32_t vo, R v, uint32_t &rs, int k) const {
    store((q + 1) >> 1, vo, 0u, v, rs, k);
  }
};

// ---------------------------------------------------------------------------
// Collision.  The reference computes f = M^-1 [ m - S (m - meq) + (1 - S/2) G ]
// with m = M F (SP:222-241).  Its relaxation vector S (SP:131) gives every
// non-conserved moment of even parity (rows 1, 2, 9-15 of M, even functions of
// e) the rate s_v and every non-conserved odd one (rows 4, 6, 8, 16-18) the
// rate s_o, and its meq (SP:209-215) is exactly M feq(1, u) outside m0.  The
// collision is therefore, direction by direction,
//   f'_q = f_q - s_v (f+_q - E+_q) - s_o (f-_q - E-_q) + F_q,
// f+ / f- the parts of f even / odd under e -> -e, E+_q = w_q (rho + 4.5 (e.u)^2
// - 1.5 u.u), E-_q = 3 w_q e.j with j = sum e f (so the conserved m0 and j
// relax by exactly nothing), and the force term split the same way:
//   F_q = w_q [ (1 - s_v/2)(-a u.F + b (e.u)(e.F)) + (s_v/2) G0 + a e.F ],
// G0 = -(3a - b)/3 u.F (a, b = 1/3, 1/9 in SP:236).  With g = f - w every E+
// term loses its w_q and rho becomes drho = rho - 1.
//
// Inputs: opposite-pair sums P[k] = g_lo + g_hi and differences D[k] = g_lo -
// g_hi (k = 1..9, lo = 2k-1), rest population g0; outputs overwrite g.
// ---------------------------------------------------------------------------
template <typename R>
struct Macro {
  R drho, jx, jy, jz, ux, uy, uz;
};

// e_q . (x, y, z) with the zero and unit components folded at compile time
// (q is a constant after unrolling; 0 * x is not 0 for every float, so the
// products are not written out).
template <typename R>
__device__ __forceinline__ R edot(int q, R x, R y, R z) {
  const int a = ex(q), b = ey(q), c = ez(q);
  R r = R(0);
  if (a != 0) r = a > 0 ? x : -x;
  if (b != 0) r = (a != 0) ? (b > 0 ? r + y : r - y) : (b > 0 ? y : -y);
  if (c != 0) r = (a != 0 || b != 0) ? (c > 0 ? r + z : r - z) : (c > 0 ? z : -z);
  return r;
}

template <typename R>
__device__ __forceinline__ void pair_sums(const R g[kQ], R P[10], R D[10]) {
#pragma unroll
  for (int k = 1; k <= 9; ++k) {
    P[k] = g[2 * k - 1] + g[2 * k];
    D[k] = g[2 * k - 1] - g[2 * k];
  }
}

// rho and v of streaming3 (SP:380-388): v = (j + F/2) / rho.  `first` drops
// the half-force shift: init's state carries v = u0 exactly (SP:165-166).
template <typename R>
__device__ __forceinline__ Macro<R> macros_pd(R g0, const R P[10], const R D[10], const Phys<R> &ph, bool first) {
  Macro<R> m;
  m.drho = g0 + ((P[1] + P[2]) + (P[3] + P[4])) + ((P[5] + P[6]) + (P[7] + P[8]) + P[9]);
  // e_lo of the pairs: 1 (1,0,0) 2 (0,1,0) 3 (0,0,1) 4 (1,1,0) 5 (1,-1,0)
  //                    6 (1,0,1) 7 (1,0,-1) 8 (0,1,1) 9 (0,1,-1)
  m.jx = D[1] + (D[4] + D[5]) + (D[6] + D[7]);
  m.jy = D[2] + (D[4] - D[5]) + (D[8] + D[9]);
  m.jz = D[3] + (D[6] - D[7]) + (D[8] - D[9]);
  const R inv = R(1) / (R(1) + m.drho);
  if (first) {
    m.ux = m.jx * inv;
    m.uy = m.jy * inv;
    m.uz = m.jz * inv;
  } else {
    m.ux = (m.jx + R(0.5) * ph.fx) * inv;
    m.uy = (m.jy + R(0.5) * ph.fy) * inv;
    m.uz = (m.jz + R(0.5) * ph.fz) * inv;
  }
  return m;
}

template <typename R, bool FORCED>
__device__ __forceinline__ void collide_impl(R g[kQ], const R P[10], const R D[10], const Macro<R> &m,
                                             const Phys<R> &ph) {
  const R sv = ph.sv, so = ph.so;
  const R hsv = R(0.5) * sv, hso = R(0.5) * so;
  const R ux = m.ux, uy = m.uy, uz = m.uz;
  const R c0 = m.drho - R(1.5) * (ux * ux + uy * uy + uz * uz);
  R K0 = R(0), K1 = R(0), Fa = R(0);
  const R Fx = ph.fx, Fy = ph.fy, Fz = ph.fz;
  if constexpr (FORCED) {
    const R a = ph.ga, b = ph.gb;
    const R uF = ux * Fx + uy * Fy + uz * Fz;
    const R G0 = -(R(3) * a - b) * (R(1) / R(3)) * uF;
    K0 = -(R(1) - hsv) * a * uF + hsv * G0;
    K1 = (R(1) - hsv) * b;
    Fa = a;
  }
  // rest direction
  {
    const R w = wq<R>(0);
    g[0] = g[0] - sv * (g[0] - w * c0);
    if constexpr (FORCED) g[0] += w * K0;
  }
#pragma unroll
  for (int k = 1; k <= 9; ++k) {
    const int q = 2 * k - 1;
    const R w = wq<R>(q);
    const R eu = edot<R>(q, ux, uy, uz);
    const R ej = edot<R>(q, m.jx, m.jy, m.jz);
    const R ap = hsv * P[k] - (sv * w) * (c0 + R(4.5) * eu * eu);  // s_v (g+ - E+)
    const R am = hso * D[k] - (so * R(3) * w) * ej;                // s_o (g- - E-)
    R lo = g[q] - ap - am, hi = g[q + 1] - ap + am;
    if constexpr (FORCED) {
      const R eF = edot<R>(q, Fx, Fy, Fz);
      const R fp = w * (K0 + K1 * eu * eF);
      const R fm = (w * Fa) * eF;
      lo += fp + fm;
      hi += fp - fm;
    }
    g[q] = lo;
    g[q + 1] = hi;
  }
}

// feq(q, rho, u) - w_q  (SP:152-158, shifted)
template <typename R>
__device__ __forceinline__ R feq_shift(int q, R rho, R ux, R uy, R uz, R uu) {
  const R eu = edot<R>(q, ux, uy, uz);
  return wq<R>(q) * (rho - R(1)) + wq<R>(q) * rho * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu);
}

// ---------------------------------------------------------------------------
// Boundary faces.  Owner of a node = the last face, in the order
// x-left, x-right, y-left, y-right, z-left, z-right, whose mode is 1 or 2 and
// which contains the node (Boundary_condition applies faces in that order,
// SP:274-370, so later faces overwrite earlier ones).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int face_owner(const lbm3d_params &p, const Geo &gm, int i, int j, int k) {
  int owner = -1;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int m = p.bc_mode[f];
    if (m != 1 && m != 2) continue;
    const int pos = f < 2 ? i : (f < 4 ? j : k);
    const int n = f < 2 ? gm.nx : (f < 4 ? gm.ny : gm.nz);
    if (pos == ((f & 1) ? n - 1 : 0)) owner = f;
  }
  return owner;
}

// index of (i,j,k) in face f's plane
__device__ __forceinline__ uint32_t plane_index(const Geo &gm, int f, int i, int j, int k) {
  return f < 2 ? (uint32_t)j * gm.nz + k : (f < 4 ? (uint32_t)i * gm.nz + k : (uint32_t)i * gm.ny + j);
}

// Boundary_condition (SP:272-370) for a face-owned node: every population is
// replaced by feq of the owner face.  Pressure (mode 1): rho_bc and the node's
// previous-step velocity, or 0 when the inward neighbour is solid (its v is 0,
// SP:390-392).  Velocity (mode 2): rho 1 and the face velocity.
template <typename R>
__device__ __forceinline__ void face_populations(R v[kQ], const KP &kp, const Geo &gm, int owner, bool inward_solid,
                                                 const R *__restrict__ face_u, uint32_t pidx) {
  const lbm3d_params &p = kp.p;
  const Phys<R> &ph = phys<R>(kp);
  R rho, ux, uy, uz;
  if (p.bc_mode[owner] == 1) {
    rho = ph.rho_bc[owner];
    if (inward_solid) {
      ux = uy = uz = R(0);
    } else {
      const R *fu = face_u + gm.face_u_off[owner];
      const uint32_t P = gm.plane[owner];
      ux = fu[pidx];
      uy = fu[P + pidx];
      uz = fu[2 * P + pidx];
    }
  } else {
    rho = R(1);
    ux = ph.u_bc[owner][0];
    uy = ph.u_bc[owner][1];
    uz = ph.u_bc[owner][2];
  }
  const R uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int q = 0; q < kQ; ++q) v[q] = feq_shift<R>(q, rho, ux, uy, uz, uu);
}

// ---------------------------------------------------------------------------
// The read stage shared by the step and macros kernels.
//
// Even t: cell x's populations sit in its own slots g[q][x].
// Odd t : population q of cell x sits in slot opp(q) of the cell at x - e_q,
//         or, when that cell is not fluid, in slot q of x itself (the
//         bounce-back x wrote on the previous step).  Equivalently, for each q
//         the address  A_q = (nb_q(x) >= 0) ? g[q][nb_q(x)] : g[opp q][x]
//         holds population opp(q); the step writes post-collision f_q back to
//         A_q, so every cell reads and writes the same set of slots and cells'
//         sets are disjoint (race free in place).  off[q] is A_q as a byte
//         offset within q's slot pair.
// ---------------------------------------------------------------------------
struct CellCtx {
  int i, j, k;
  bool have_coords;
};

// Row words for the odd-step neighbour lookup: the table in memory, or, when
// win != nullptr, the LDS window of the cell's plane segment (3 planes x JW
// rows x W words around rows (i0, j0 ..); a tile that spans two segments
// stages one window for each, and each lane points at its own).
struct RowView {
  const Rows *rb;
  const uint4 *win;
  uint32_t JW;
  int i0, j0;
  uint32_t c0;  // dense index of cell (i0, j0, 0) (row_window's one-segment path)
  int32_t cell; // >= 0: the cell's dense index, loaded ahead (else read from cells[])
  int jrel, k;  // jrel >= 0: the cell's row (from j0) in the window and its k, found without cells[]
};

// Odd step: byte offsets off[q] of the addresses A_q within q's slot pair, and
// the cell's coordinates.
template <typename T>
__device__ __forceinline__ void odd_offsets(const Slots<T> &sl, const Geo &gm, const int32_t *__restrict__ cells,
                                            const RowView &rv, uint32_t x, uint32_t off[kQ], CellCtx &cc) {
  const uint32_t xb = x * (uint32_t)sizeof(T);
  cc.have_coords = true;
  int32_t nb[18];
  const uint32_t dense = rv.cell >= 0 ? (uint32_t)rv.cell : (rv.win && rv.jrel >= 0 ? 0u : (uint32_t)cells[x]);
  if (rv.win) {
    // the window's cells all lie in plane i0, rows j0 .. j0+J
    uint32_t jrel;
    if (rv.jrel >= 0) {
      jrel = (uint32_t)rv.jrel;
      cc.k = rv.k;
    } else {
      const uint32_t d = dense - rv.c0;
      jrel = fdiv(d, gm.m_nz, gm.s_nz);
      cc.k = (int)(d - jrel * (uint32_t)gm.nz);
    }
    cc.i = rv.i0;
    cc.j = rv.j0 + (int)jrel;
    const WindowRows src(rv.win, gm.W, rv.JW, (int)jrel, cc.k);
    neighbor_ids(src, gm, cc.k, nb);
  } else {
    coords(gm, dense, cc.i, cc.j, cc.k);
    const TableRows src(*rv.rb, gm, cc.i, cc.j);
    neighbor_ids(src, gm, cc.k, nb);
  }
  off[0] = xb;
#pragma unroll
  for (int q = 1; q < kQ; ++q) {
    const int32_t c = nb[q - 1];
    const uint32_t self_slot = (q & 1) ? sl.Sb : 0u;  // slot opp(q) within the pair
    const uint32_t nb_slot = (q & 1) ? 0u : sl.Sb;    // slot q within the pair
    off[q] = c >= 0 ? nb_slot + (uint32_t)c * (uint32_t)sizeof(T) : self_slot + xb;
  }
}

template <typename T, bool ODD>
__device__ __forceinline__ void read_cell(const Slots<T> &sl, const Geo &gm, const int32_t *__restrict__ cells,
                                          const RowView &rv, uint32_t x, typename St<T>::R v[kQ], uint32_t off[kQ],
                                          CellCtx &cc) {
  const uint32_t xb = x * (uint32_t)sizeof(T);
  if constexpr (!ODD) {
    cc.have_coords = false;
#pragma unroll
    for (int q = 0; q < kQ; ++q) v[q] = sl.ld_slot(q, xb);
  } else {
    odd_offsets<T>(sl, gm, cells, rv, x, off, cc);
    v[0] = sl.ld_slot(0, xb);
#pragma unroll
    for (int q = 1; q < kQ; ++q) v[opp(q)] = sl.ld_pair(q, off[q]);
  }
}

// 2-bit face record of cell x: bit 0 face-owned, bit 1 owner's inward neighbour solid
__device__ __forceinline__ uint32_t face_record(const uint32_t *__restrict__ face_bits, uint32_t x) {
  return (face_bits[x >> 4] >> ((x & 15u) * 2u)) & 3u;
}

// One fluid cell's time step: read (with the previous step's face overwrite),
// macros, pressure-face state, collision, write back in place.
// The cell's update between the read and the write: the previous step's face
// overwrite, macros, the pressure-face state, collision (in place on v).
template <typename R, bool FORCED>
__device__ __forceinline__ void collide_cell(const KP &kp, const Geo &gm, const int32_t *__restrict__ cells,
                                             const uint32_t *__restrict__ face_bits, R *__restrict__ face_u,
                                             uint32_t x, int first, R v[kQ], CellCtx &cc) {
  const lbm3d_params &p = kp.p;
  const Phys<R> &ph = phys<R>(kp);
  const uint32_t fr = face_record(face_bits, x);
  int owner = -1;
  uint32_t pidx = 0;
  if (fr & 1u) {
    if (!cc.have_coords) coords(gm, (uint32_t)cells[x], cc.i, cc.j, cc.k);
    owner = face_owner(p, gm, cc.i, cc.j, cc.k);
    pidx = plane_index(gm, owner, cc.i, cc.j, cc.k);
    if (!first) face_populations<R>(v, kp, gm, owner, (fr & 2u) != 0, face_u, pidx);
  }
  R P[10], D[10];
  pair_sums<R>(v, P, D);
  const Macro<R> m = macros_pd<R>(v[0], P, D, ph, first != 0);
  if (owner >= 0 && p.bc_mode[owner] == 1 && !(fr & 2u)) {
    // the pressure face of the next step uses this step's v (SP:281)
    R *fu = face_u + gm.face_u_off[owner];
    const uint32_t PL = gm.plane[owner];
    fu[pidx] = m.ux;
    fu[PL + pidx] = m.uy;
    fu[2 * PL + pidx] = m.uz;
  }
  collide_impl<R, FORCED>(v, P, D, m, ph);
}

// One fluid cell's time step: read (with the previous step's face overwrite),
// macros, pressure-face state, collision, write back in place.  id0: the
// first compact id of the cell's grid (a batch sample's cell_begin, else 0).
template <typename T, bool ODD, bool FORCED>
__device__ __forceinline__ void step_cell(const KP &kp, const Geo &gm, const Slots<T> &sl, const RowView &rv,
                                          const int32_t *__restrict__ cells, const uint32_t *__restrict__ face_bits,
                                          typename St<T>::R *__restrict__ face_u, uint32_t x, int first,
                                          uint32_t id0 = 0u) {
  using R = typename St<T>::R;
  R v[kQ];
  uint32_t off[kQ];
  CellCtx cc;
  read_cell<T, ODD>(sl, gm, cells, rv, x, v, off, cc);
  collide_cell<R, FORCED>(kp, gm, cells, face_bits, face_u, x, first, v, cc);
  uint32_t rs = sl.stream(x - id0);
  if constexpr (!ODD) {
    const uint32_t xb = x * (uint32_t)sizeof(T);
#pragma unroll
    for (int q = 0; q < kQ; ++q) sl.st_slot(opp(q), xb, v[q], rs, q);
  } else {
    sl.st_slot(0, off[0], v[0], rs, 0);
#pragma unroll
    for (int q = 1; q < kQ; ++q) sl.st_pair(q, off[q], v[q], rs, q);
  }
}

// rho and v of one fluid cell as streaming3 leaves them after the last step
template <typename T, bool ODD>
__device__ __forceinline__ void macros_cell(const KP &kp, const Geo &gm, const Slots<T> &sl, const RowView &rv,
                                            const int32_t *__restrict__ cells, const uint32_t *__restrict__ face_bits,
                                            const typename St<T>::R *__restrict__ face_u, uint32_t x, int first,
                                            typename St<T>::R *__restrict__ rho_out,
                                            typename St<T>::R *__restrict__ v_out, uint32_t N) {
  using R = typename St<T>::R;
  R v[kQ];
  uint32_t off[kQ];
  CellCtx cc;
  read_cell<T, ODD>(sl, gm, cells, rv, x, v, off, cc);
  const uint32_t fr = face_record(face_bits, x);
  if ((fr & 1u) && !first) {
    if (!cc.have_coords) coords(gm, (uint32_t)cells[x], cc.i, cc.j, cc.k);
    const int owner = face_owner(kp.p, gm, cc.i, cc.j, cc.k);
    face_populations<R>(v, kp, gm, owner, (fr & 2u) != 0, face_u, plane_index(gm, owner, cc.i, cc.j, cc.k));
  }
  R P[10], D[10];
  pair_sums<R>(v, P, D);
  const Macro<R> m = macros_pd<R>(v[0], P, D, phys<R>(kp), first != 0);
  rho_out[x] = R(1) + m.drho;
  v_out[x] = m.ux;
  v_out[(size_t)N + x] = m.uy;
  v_out[2 * (size_t)N + x] = m.uz;
}

// init (SP:160-170) of one fluid cell: equilibrium at (rho0, u0), default rest
template <typename T>
__device__ __forceinline__ void init_cell(const KP &kp, const Geo &gm, const Slots<T> &sl,
                                          const int32_t *__restrict__ cells, typename St<T>::R *__restrict__ face_u,
                                          const float *__restrict__ rho0, const float *__restrict__ u0, uint32_t x,
                                          uint32_t N) {
  using R = typename St<T>::R;
  const lbm3d_params &p = kp.p;
  R rho = R(1), ux = R(0), uy = R(0), uz = R(0);
  if (rho0) rho = (R)rho0[x];
  if (u0) {
    ux = (R)u0[x];
    uy = (R)u0[(size_t)N + x];
    uz = (R)u0[2 * (size_t)N + x];
  }
  const R uu = ux * ux + uy * uy + uz * uz;
  const uint32_t xb = x * (uint32_t)sizeof(T);
  uint32_t rs = 0u;  // init rounds to nearest (sl has no seed)
#pragma unroll
  for (int q = 0; q < kQ; ++q) sl.st_slot(q, xb, feq_shift<R>(q, rho, ux, uy, uz, uu), rs, q);
  int i, j, k;
  coords(gm, (uint32_t)cells[x], i, j, k);
  const int owner = face_owner(p, gm, i, j, k);
  if (owner >= 0 && p.bc_mode[owner] == 1) {
    R *fu = face_u + gm.face_u_off[owner];
    const uint32_t PL = gm.plane[owner], pi = plane_index(gm, owner, i, j, k);
    fu[pi] = ux;
    fu[PL + pi] = uy;
    fu[2 * PL + pi] = uz;
  }
}

// A workgroup's B cells x0 .. x0+B-1 (x0 < n) usually lie in a few
// consecutive rows (i0, j0 .. j0+J) of one plane: stage the rows i0-1 .. i0+1 x
// j0-1 .. j0+J+1 of the row table in LDS so each cell's 9 row words are LDS
// reads, not global loads.  Falls back to the table (rv.win == nullptr) when
// the cells span planes or the window would not fit.  Every thread of the
// workgroup must call it (it synchronises).
template <int B = kBlock>
__device__ __forceinline__ RowView row_window(const Geo &gm, const int32_t *__restrict__ cells,
                                              const uint4 *__restrict__ rows, const Rows &rb, uint32_t x0,
                                              uint32_t n, uint4 *win) {
  const uint32_t xl = x0 + B - 1 < n ? x0 + B - 1 : n - 1;
  int i0, j0, k0, i1, j1, k1;
  coords(gm, (uint32_t)cells[x0], i0, j0, k0);
  coords(gm, (uint32_t)cells[xl], i1, j1, k1);
  const uint32_t JW = (uint32_t)(j1 - j0) + 3u;
  // one plane's rows of one band (a block of a small grid can span a whole
  // cycle of a band's planes and end in the plane it started in)
  if (i0 == i1 && (uint32_t)j0 / gm.band == (uint32_t)j1 / gm.band && JW * 3u * gm.W <= (uint32_t)kWinWords) {
    const uint32_t per_plane = JW * gm.W, cnt = 3u * per_plane;
    for (uint32_t e = threadIdx.x; e < cnt; e += B) {
      const uint32_t di = e / per_plane, r = e - di * per_plane;
      const uint32_t dj = r / gm.W, w = r - dj * gm.W;
      const uint32_t row =
          (uint32_t)wrap(i0 + (int)di - 1, gm.nx) * (uint32_t)gm.ny + (uint32_t)wrap(j0 + (int)dj - 1, gm.ny);
      win[e] = rows[row * gm.W + w];
    }
    __syncthreads();
    return RowView{&rb, win, JW, i0, j0, ((uint32_t)i0 * (uint32_t)gm.ny + (uint32_t)j0) * (uint32_t)gm.nz, -1, -1, 0};
  }
  return RowView{&rb, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
}

// Tile descriptors (LBM3D_FLAG_TILES), one per block of kBlock cells.  The
// block's cells lie in
//   one plane's segment of the list: {i0, j0, J, 0}, rows (i0, j0 .. j0+J);
//   two consecutive segments: {i0, j0, JA | JB << 8 | kTwoSeg, iB | jB << 16},
//     rows (i0, j0 .. j0+JA) then (iB, jB .. jB+JB), the second segment being
//     the plane after i0 in the same band or the first plane of the next band;
//   else {i0, j0, kNoTile, 0}: the odd step reads the cell list and the table.
// A descriptor's rows, with the rows and planes around them, fit kWinWords
// words of LDS and its middle rows' words fit one wave (at most 64).  The odd
// step stages that window straight from the descriptor and finds every cell's
// place from the window's words, without reading the cell list.
constexpr uint32_t kNoTile = 0xFFFFFFFFu;
constexpr uint32_t kTwoSeg = 1u << 16;

// A descriptor's segments: segment A at win[0 .. offB), B (when nB > 0) after it;
// each is 3 planes x JW rows x W words, its middle rows' words at e (nA / nB of them).
struct TileSegs {
  int iA, jA, iB, jB;
  uint32_t JWA, JWB, nA, nB, offB;
};

__device__ __forceinline__ TileSegs tile_segs(const Geo &gm, const uint4 td) {
  const bool two = (td.z & kTwoSeg) != 0u;
  const uint32_t JA = two ? (td.z & 0xFFu) : td.z, JB = (td.z >> 8) & 0xFFu;
  TileSegs t;
  t.iA = (int)td.x;
  t.jA = (int)td.y;
  t.iB = (int)(td.w & 0xFFFFu);
  t.jB = (int)(td.w >> 16);
  t.JWA = JA + 3u;
  t.nA = (JA + 1u) * gm.W;
  t.JWB = two ? JB + 3u : 0u;
  t.nB = two ? (JB + 1u) * gm.W : 0u;
  t.offB = 3u * t.JWA * gm.W;
  return t;
}

// Stage rows i-1 .. i+1 x j-1 .. j+JW-2 of each segment (periodic wrap) in
// LDS.  Every thread of the workgroup must call it (it synchronises).
template <int B = kBlock>
__device__ __forceinline__ void tile_stage(const Geo &gm, const uint4 *__restrict__ rows, const TileSegs &t,
                                           uint4 *win) {
  const uint32_t cnt = t.offB + 3u * t.JWB * gm.W;
  for (uint32_t e = threadIdx.x; e < cnt; e += B) {
    const bool b = e >= t.offB;
    const uint32_t el = b ? e - t.offB : e, per_plane = (b ? t.JWB : t.JWA) * gm.W;
    const uint32_t di = el / per_plane, r = el - di * per_plane;
    const uint32_t dj = r / gm.W, w = r - dj * gm.W;
    const uint32_t row = (uint32_t)wrap((b ? t.iB : t.iA) + (int)di - 1, gm.nx) * (uint32_t)gm.ny +
                         (uint32_t)wrap((b ? t.jB : t.jA) + (int)dj - 1, gm.ny);
    win[e] = rows[row * gm.W + w];
  }
  __syncthreads();
}

// position of the r-th (from 0) set bit of m (r < popcount(m))
__device__ __forceinline__ uint32_t select_bit(uint32_t m, uint32_t r) {
  uint32_t pos = 0u, c;
  c = __popc(m & 0xFFFFu);
  if (r >= c) { r -= c; m >>= 16; pos += 16u; }
  c = __popc(m & 0xFFu);
  if (r >= c) { r -= c; m >>= 8; pos += 8u; }
  c = __popc(m & 0xFu);
  if (r >= c) { r -= c; m >>= 4; pos += 4u; }
  c = __popc(m & 0x3u);
  if (r >= c) { r -= c; m >>= 2; pos += 2u; }
  return pos + ((m & 1u) <= r ? 1u : 0u);
}

// Row (jrel, from its segment's first row) and k of cell x, and the window of
// its segment, from the segments' middle-row words (nA + nB <= 64, in list
// order): lane l holds word l; an exclusive suffix minimum gives empty words
// the first id of the next word with a fluid cell, and a binary search over
// the lanes finds the last word whose first id is <= x.  Every lane of the
// wave must call it.
__device__ __forceinline__ void tile_place(const Geo &gm, const uint4 *win, const TileSegs &t, uint32_t x,
                                           RowView &rv) {
  const uint32_t W = gm.W, nw = t.nA + t.nB;
  const uint32_t eA = (t.JWA + 1u) * W, eB = t.offB + (t.JWB + 1u) * W;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t key = 0x7FFFFFFFu;
  if (lane < nw) {
    const uint4 wd = win[lane < t.nA ? eA + lane : eB + (lane - t.nA)];
    if (wd.x) key = wd.y;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = (uint32_t)__shfl_down((int)key, off);
    if (lane + (uint32_t)off < 64u) key = min(key, o);
  }
  uint32_t pos = 0u;
#pragma unroll
  for (uint32_t step = 32u; step > 0u; step >>= 1) {
    const uint32_t cand = pos + step;
    const uint32_t kc = (uint32_t)__shfl((int)key, (int)(cand < 64u ? cand : 63u));
    if (cand < nw && kc <= x) pos = cand;
  }
  const bool segb = pos >= t.nA;
  const uint32_t p = segb ? pos - t.nA : pos;
  const uint4 hw = win[segb ? eB + p : eA + p];
  rv.k = (int)(32u * (p % W) + select_bit(hw.x, x - hw.y));
  rv.jrel = (int)(p / W);
  rv.win = segb ? win + t.offB : win;
  rv.JW = segb ? t.JWB : t.JWA;
  rv.i0 = segb ? t.iB : t.iA;
  rv.j0 = segb ? t.jB : t.jA;
}

// One tile (kBlock consecutive fluid cells from lb * kBlock) of a time step.
// win: the workgroup's LDS row window (odd steps).  Every thread of the
// workgroup must call it (the window is staged behind a barrier).
template <typename T, bool ODD, bool FORCED>
__device__ __forceinline__ void step_tile(const KP &kp, T *__restrict__ g, const int32_t *__restrict__ cells,
                                          const uint4 *__restrict__ rows, const uint32_t *__restrict__ face_bits,
                                          typename St<T>::R *__restrict__ face_u, int first, uint32_t seed,
                                          uint32_t lb, uint4 *win) {
  const uint32_t n = (uint32_t)kp.p.n_cells;
  const uint32_t x0 = lb * kBlock;
  const uint32_t x = x0 + threadIdx.x;
  const Slots<T> sl(g, kp.S, seed);
  const Rows rb(rows, kp.geo.n_row_words);
  RowView rv{&rb, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
  if constexpr (ODD && LBM3D_ROW_WINDOW) {
    if (kp.p.flags & LBM3D_FLAG_TILES) {
      const uint4 td = rows[kp.geo.n_row_words + lb];  // this block's descriptor
      if (td.z != kNoTile) {
        const TileSegs ts = tile_segs(kp.geo, td);
        tile_stage(kp.geo, rows, ts, win);
        tile_place(kp.geo, win, ts, x, rv);
      } else {
        rv.cell = cells[x < n ? x : n - 1];  // table path
      }
    } else {
      // the cell's dense index is loaded before the window is staged, so the
      // two loads overlap instead of following each other
      const int32_t own = cells[x < n ? x : n - 1];
      rv = row_window(kp.geo, cells, rows, rb, x0, n, win);
      rv.cell = own;
    }
  }
  if (x >= n) return;
  step_cell<T, ODD, FORCED>(kp, kp.geo, sl, rv, cells, face_bits, face_u, x, first);
}

// minimum waves per SIMD the step kernel of storage T is compiled for
template <typename T> struct StepWaves { static constexpr int v = LBM3D_STEP_MINB; };
template <> struct StepWaves<__half> { static constexpr int v = LBM3D_STEP_WAVES16; };
template <> struct StepWaves<__hip_bfloat16> { static constexpr int v = LBM3D_STEP_WAVES16; };

// One time step over the blocks lb_begin .. lb_begin + nb - 1 of the fluid
// list.
template <typename T, bool ODD, bool FORCED>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kBlock), amdgpu_waves_per_eu(StepWaves<T>::v)))
    k_step(const KP kp, T *__restrict__ g, const int32_t *__restrict__ cells, const uint4 *__restrict__ rows,
           const uint32_t *__restrict__ face_bits, typename St<T>::R *__restrict__ face_u, int first,
           uint32_t lb_begin, uint32_t nb) {
  const uint32_t lb = logical_block<ODD ? LBM3D_XCD_CHUNK_ODD : LBM3D_XCD_CHUNK_EVEN,
                                    !ODD && LBM3D_REVERSE_EVEN>(nb);
  if (lb >= nb) return;  // padding blocks of the grid
  __shared__ uint4 win[ODD && LBM3D_ROW_WINDOW ? kWinWords : 1];
  step_tile<T, ODD, FORCED>(kp, g, cells, rows, face_bits, face_u, first, kp.sr_seed, lb_begin + lb, win);
}

template <typename T, bool ODD>
__global__ void __launch_bounds__(kBlock)
    k_macros(const KP kp, const T *__restrict__ g, const int32_t *__restrict__ cells,
             const uint4 *__restrict__ rows, const uint32_t *__restrict__ face_bits,
             const typename St<T>::R *__restrict__ face_u, int first, typename St<T>::R *__restrict__ rho_out,
             typename St<T>::R *__restrict__ v_out) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)kp.p.n_cells;
  if (x >= N) return;
  const Slots<T> sl(g, kp.S);
  const Rows rb(rows, kp.geo.n_row_words);
  const RowView rv{&rb, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
  macros_cell<T, ODD>(kp, kp.geo, sl, rv, cells, face_bits, face_u, x, first, rho_out, v_out, N);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_init(const KP kp, T *__restrict__ g, const int32_t *__restrict__ cells, typename St<T>::R *__restrict__ face_u,
           const float *__restrict__ rho0, const float *__restrict__ u0) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)kp.p.n_cells;
  if (x >= N) return;
  const Slots<T> sl(g, kp.S);
  init_cell<T>(kp, kp.geo, sl, cells, face_u, rho0, u0, x, N);
}

// ---------------------------------------------------------------------------
// Batches.  Samples of one batch share the physical parameters and face modes;
// each has its own grid.  Their fluid lists are laid end to end in one slot
// range, each starting on a kChunk-cell boundary, and every row word holds
// batch-wide compact ids, so the neighbour lookup and the population accesses
// are those of a single grid.  chunk_sample[c] names the sample owning cells
// c*kChunk .. c*kChunk + kChunk-1; a workgroup of B | kChunk cells therefore
// lies in one sample and reads that sample's descriptor with scalar loads.
// ---------------------------------------------------------------------------
constexpr uint32_t kChunk = LBM3D_BATCH_ALIGN;
static_assert(kChunk % kBlock == 0, "a block of the batch lies in one sample");

struct SampleDesc {
  Geo geo;
  uint32_t cell_begin, cell_end;    // the sample's fluid cells: compact ids [cell_begin, cell_end)
  uint32_t chunk_begin, chunk_end;  // its kChunk-cell chunks
  uint64_t rows_begin;              // its first row word
  uint64_t tiles_begin;             // row-buffer index of its first block's tile descriptor (LBM3D_FLAG_TILES)
};

struct Batch {
  const SampleDesc *desc;
  const int32_t *chunk_sample;
  const uint8_t *active;            // per sample: 0 = skip (retired), or NULL = all active
};

// Persistent launch: gridDim.x workgroups walk the n_tiles tiles of B cells,
// forwards or backwards.  With C > 0 (a grid of 8 * per_xcd workgroups) each
// XCD walks runs of C consecutive tiles (xcd.h); odd steps stage each tile's
// row window in LDS, as k_step does.
template <typename T, bool ODD, bool FORCED, int B>
__global__ void __launch_bounds__(B)
    k_step_batch(const KP kp, const Batch bt, T *__restrict__ g, const int32_t *__restrict__ cells,
                 const uint4 *__restrict__ rows, const uint32_t *__restrict__ face_bits,
                 typename St<T>::R *__restrict__ face_u, int first, uint32_t tile0, uint32_t n_tiles, int reverse) {
  constexpr uint32_t C = ODD ? LBM3D_BATCH_XCD_CHUNK_ODD : LBM3D_BATCH_XCD_CHUNK_EVEN;
  const Slots<T> sl(g, kp.S, kp.sr_seed);
  const uint32_t per_xcd = gridDim.x >> 3;
  for (uint32_t i = 0;; ++i) {
    uint32_t it;
    if constexpr (C > 0) {
      it = xcd_block(C, blockIdx.x & 7u, (blockIdx.x >> 3) + i * per_xcd, n_tiles);
    } else {
      it = blockIdx.x + i * gridDim.x;
    }
    if (it >= n_tiles) break;  // every condition below is uniform in the workgroup
    const uint32_t x0 = (tile0 + (reverse ? n_tiles - 1u - it : it)) * (uint32_t)B;
    const int s = bt.chunk_sample[x0 / kChunk];
    if (bt.active && !bt.active[s]) continue;
    const SampleDesc &d = bt.desc[s];
    if (x0 >= d.cell_end) continue;
    const uint32_t x = x0 + threadIdx.x;
    const Rows rb(rows + d.rows_begin, d.geo.n_row_words);
    RowView rv{&rb, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
    if constexpr (ODD && LBM3D_ROW_WINDOW) {
      __shared__ uint4 win[kWinWords];
      const int32_t own = cells[x < d.cell_end ? x : d.cell_end - 1];
      __syncthreads();  // the previous tile's window reads are done
      rv = row_window<B>(d.geo, cells, rows + d.rows_begin, rb, x0, d.cell_end, win);
      rv.cell = own;
    }
    if (x < d.cell_end) step_cell<T, ODD, FORCED>(kp, d.geo, sl, rv, cells, face_bits, face_u, x, first, d.cell_begin);
  }
}

// One tile per workgroup (the launch shapes with no persistent grid): the
// same per-tile work as k_step_batch without its loop, which keeps the odd
// step's registers at the single-grid kernel's count.
template <typename T, bool ODD, bool FORCED, int B>
__global__ void __launch_bounds__(B)
    k_step_batch1(const KP kp, const Batch bt, T *__restrict__ g, const int32_t *__restrict__ cells,
                  const uint4 *__restrict__ rows, const uint32_t *__restrict__ face_bits,
                  typename St<T>::R *__restrict__ face_u, int first, uint32_t tile0, uint32_t n_tiles,
                  int reverse) {
  constexpr uint32_t C = ODD ? LBM3D_BATCH_XCD_CHUNK_ODD : LBM3D_BATCH_XCD_CHUNK_EVEN;
  uint32_t it = blockIdx.x;
  if constexpr (C > 0) it = xcd_block(C, blockIdx.x & 7u, blockIdx.x >> 3, n_tiles);
  if (it >= n_tiles) return;  // padding block of the grid
  const uint32_t x0 = (tile0 + (reverse ? n_tiles - 1u - it : it)) * (uint32_t)B;
  const uint32_t x = x0 + threadIdx.x;
  const Slots<T> sl(g, kp.S, kp.sr_seed);
  if constexpr (!ODD) {
    if (!bt.active) {
      // With every sample active, an even step issues its population loads
      // (the cell's own slots, which exist for every id of the batch, padding
      // included) before the sample lookup, instead of after two dependent
      // round trips for the chunk's sample and its descriptor.
      using R = typename St<T>::R;
      R v[kQ];
      uint32_t off[kQ];
      CellCtx cc;
      const RowView none{nullptr, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
      read_cell<T, false>(sl, kp.geo, cells, none, x, v, off, cc);
      const SampleDesc &d = bt.desc[bt.chunk_sample[x0 / kChunk]];
      if (x >= d.cell_end) return;
      collide_cell<R, FORCED>(kp, d.geo, cells, face_bits, face_u, x, first, v, cc);
      const uint32_t xb = x * (uint32_t)sizeof(T);
      uint32_t rs = sl.stream(x - d.cell_begin);
#pragma unroll
      for (int q = 0; q < kQ; ++q) sl.st_slot(opp(q), xb, v[q], rs, q);
      return;
    }
  }
  const int s = bt.chunk_sample[x0 / kChunk];
  if (bt.active && !bt.active[s]) return;
  const SampleDesc &d = bt.desc[s];
  if (x0 >= d.cell_end) return;
  const Rows rb(rows + d.rows_begin, d.geo.n_row_words);
  RowView rv{&rb, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
  if constexpr (ODD && LBM3D_ROW_WINDOW) {
    __shared__ uint4 win[kWinWords];
    bool placed = false;
    if constexpr (B == kBlock) {
      if (kp.p.flags & LBM3D_FLAG_TILES) {  // as k_step: the window from this block's descriptor
        const uint4 td = rows[d.tiles_begin + (x0 - d.cell_begin) / kBlock];
        if (td.z != kNoTile) {
          const TileSegs ts = tile_segs(d.geo, td);
          tile_stage(d.geo, rows + d.rows_begin, ts, win);
          tile_place(d.geo, win, ts, x, rv);
        } else {
          rv.cell = cells[x < d.cell_end ? x : d.cell_end - 1];  // table path
        }
        placed = true;
      }
    }
    if (!placed) {
      const int32_t own = cells[x < d.cell_end ? x : d.cell_end - 1];
      rv = row_window<B>(d.geo, cells, rows + d.rows_begin, rb, x0, d.cell_end, win);
      rv.cell = own;
    }
  }
  if (x < d.cell_end) step_cell<T, ODD, FORCED>(kp, d.geo, sl, rv, cells, face_bits, face_u, x, first, d.cell_begin);
}

template <typename T, bool ODD>
__global__ void __launch_bounds__(kBlock)
    k_macros_batch(const KP kp, const Batch bt, const T *__restrict__ g, const int32_t *__restrict__ cells,
                   const uint4 *__restrict__ rows, const uint32_t *__restrict__ face_bits,
                   const typename St<T>::R *__restrict__ face_u, int first, typename St<T>::R *__restrict__ rho_out,
                   typename St<T>::R *__restrict__ v_out) {
  const uint32_t x0 = blockIdx.x * kBlock;
  const SampleDesc &d = bt.desc[bt.chunk_sample[x0 / kChunk]];
  const uint32_t x = x0 + threadIdx.x;
  if (x >= d.cell_end) return;
  const Slots<T> sl(g, kp.S);
  const Rows rb(rows + d.rows_begin, d.geo.n_row_words);
  const RowView rv{&rb, nullptr, 0u, 0, 0, 0u, -1, -1, 0};
  macros_cell<T, ODD>(kp, d.geo, sl, rv, cells, face_bits, face_u, x, first, rho_out, v_out, (uint32_t)kp.p.n_cells);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_init_batch(const KP kp, const Batch bt, T *__restrict__ g, const int32_t *__restrict__ cells,
                 typename St<T>::R *__restrict__ face_u, const float *__restrict__ rho0,
                 const float *__restrict__ u0) {
  const uint32_t x0 = blockIdx.x * kBlock;
  const SampleDesc &d = bt.desc[bt.chunk_sample[x0 / kChunk]];
  const uint32_t x = x0 + threadIdx.x;
  if (x >= d.cell_end) return;
  const Slots<T> sl(g, kp.S);
  init_cell<T>(kp, d.geo, sl, cells, face_u, rho0, u0, x, (uint32_t)kp.p.n_cells);
}

// ---------------------------------------------------------------------------
// Compaction: tiles of kTile positions (kItems per thread) of the list order;
// per tile the count of fluid cells, an exclusive scan over tiles, then a
// block-level scan inside each tile to number the cells in list order.
//
// List order: the grid rows (i, j) in bands of R consecutive j (the last band
// may be narrower); band after band, and inside a band plane after plane
// (i), row after row (j), k fastest.  R >= ny is C order.  Every row's cells
// stay consecutive and in k order, which is all the row words assume; a
// cell's neighbours in the next and previous planes are then R rows of cells
// away in the list instead of a whole plane.
// ---------------------------------------------------------------------------
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;

struct ListOrder {
  uint32_t nx, ny, nz, R;  // R: band rows, clamped to [1, ny]
};

ListOrder list_order(const lbm3d_params &p) {
  const int32_t R = p.band_rows <= 0 || p.band_rows >= p.ny ? p.ny : p.band_rows;
  return ListOrder{(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz, (uint32_t)R};
}

// dense C-order index of list position pos (pos < nx ny nz < 2^31)
__device__ __forceinline__ uint32_t list_dense(const ListOrder &o, uint32_t pos) {
  if (o.R == o.ny) return pos;
  const uint32_t row = pos / o.nz, k = pos - row * o.nz;
  const uint32_t band_rows = o.nx * o.R;  // rows of a full band
  const uint32_t jb = row / band_rows;
  const uint32_t Rb = min(o.R, o.ny - jb * o.R);
  const uint32_t r = row - jb * band_rows;
  const uint32_t i = r / Rb, j = jb * o.R + (r - i * Rb);
  return (i * o.ny + j) * o.nz + k;
}

__global__ void __launch_bounds__(kBlock) k_compact_count(const int8_t *__restrict__ solid, const ListOrder o,
                                                          int64_t n_dense, int32_t *__restrict__ tile_cnt) {
  __shared__ int s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int cnt = 0;
  for (int it = 0; it < kItems; ++it) {
    const int64_t c = base + (int64_t)threadIdx.x * kItems + it;
    if (c >= n_dense) break;
    cnt += solid[list_dense(o, (uint32_t)c)] == 0;  // fluid iff solid == 0 (SP:225)
  }
  if (cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = s_cnt;
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

// exclusive scan over tiles (single block, sequential chunks), total to count[0]
__global__ void __launch_bounds__(kBlock) k_compact_scan(int32_t *__restrict__ tile_cnt, int64_t n_tiles,
                                                         int32_t *__restrict__ count) {
  __shared__ int s_warp[kBlock / 64];
  __shared__ int s_tot;
  int carry = 0;
  for (int64_t start = 0; start < n_tiles; start += kBlock) {
    const int64_t t = start + threadIdx.x;
    const int v = t < n_tiles ? tile_cnt[t] : 0;
    const int sc = block_inclusive_scan(v, s_warp);
    if (t < n_tiles) tile_cnt[t] = carry + sc - v;
    if (threadIdx.x == kBlock - 1) s_tot = sc;
    __syncthreads();
    carry += s_tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) count[0] = carry;
}

__global__ void __launch_bounds__(kBlock) k_compact_fill(const int8_t *__restrict__ solid, const ListOrder lo,
                                                         int64_t n_dense, const int32_t *__restrict__ tile_off,
                                                         int32_t *__restrict__ d2c, int32_t *__restrict__ cells) {
  __shared__ int s_warp[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  bool fl[kItems];
  uint32_t dn[kItems];
  int cnt = 0;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int64_t c = base + (int64_t)threadIdx.x * kItems + it;
    dn[it] = c < n_dense ? list_dense(lo, (uint32_t)c) : 0u;
    fl[it] = c < n_dense && solid[dn[it]] == 0;
    cnt += fl[it];
  }
  int o = block_inclusive_scan(cnt, s_warp) - cnt + tile_off[blockIdx.x];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int64_t c = base + (int64_t)threadIdx.x * kItems + it;
    if (c >= n_dense) break;
    const int32_t id = fl[it] ? o++ : -1;
    if (d2c) d2c[dn[it]] = id;
    if (id >= 0) cells[id] = (int32_t)dn[it];
  }
}

// rows[(i*ny + j)*W + w] = {fluid mask of cells 32w .. 32w+31, compact id of the
// first of them, id of the cell before 32w, id of the cell after 32w+31}, every
// id shifted by id_offset (the sample's first compact id inside a batch)
__global__ void __launch_bounds__(kBlock) k_build_rows(const Geo gm, const int8_t *__restrict__ solid,
                                                       const int32_t *__restrict__ d2c, int32_t id_offset,
                                                       uint4 *__restrict__ rows) {
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= gm.n_row_words) return;
  const uint64_t r = t / gm.W;
  const uint32_t w = (uint32_t)(t % gm.W);
  const int nz = gm.nz;
  const int64_t row0 = (int64_t)r * nz;
  const int64_t c0 = row0 + 32 * w;
  uint32_t mask = 0;
  int32_t base = 0;
  bool found = false;
  for (int b = 0; b < 32; ++b) {
    if ((int)(32 * w) + b >= nz) break;
    if (solid[c0 + b] == 0) {
      mask |= 1u << b;
      if (!found) {
        base = d2c[c0 + b] + id_offset;
        found = true;
      }
    }
  }
  const int left = w == 0 ? nz - 1 : (int)(32 * w) - 1;            // wrap at k = 0
  const int right = (w == gm.W - 1) ? 0 : (int)(32 * w) + 32;      // wrap after k = nz-1
  const int32_t il = d2c[row0 + left], ir = d2c[row0 + right];
  rows[t] = make_uint4(mask, (uint32_t)base, (uint32_t)(il >= 0 ? il + id_offset : -1),
                       (uint32_t)(ir >= 0 ? ir + id_offset : -1));
}

// tile descriptors, one thread per block of kBlock cells
__global__ void __launch_bounds__(kBlock) k_build_tiles(const Geo gm, const int32_t *__restrict__ cells, uint32_t n,
                                                        uint4 *__restrict__ tiles) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t x0 = t * (uint32_t)kBlock;
  if (x0 >= n) return;
  const uint32_t xl = x0 + kBlock - 1 < n ? x0 + kBlock - 1 : n - 1;
  int i0, j0, k0, i1, j1, k1;
  coords(gm, (uint32_t)cells[x0], i0, j0, k0);
  coords(gm, (uint32_t)cells[xl], i1, j1, k1);
  const uint32_t bandA = (uint32_t)j0 / gm.band, bandL = (uint32_t)j1 / gm.band;
  // one segment: the block's first and last cells are in the same plane and
  // the same band of the list order, so every cell between lies in rows j0 .. j1
  if (i0 == i1 && bandA == bandL) {
    const uint32_t J = (uint32_t)(j1 - j0);
    const bool ok = 3u * (J + 3u) * gm.W <= (uint32_t)kWinWords && (J + 1u) * gm.W <= 64u;
    tiles[t] = make_uint4((uint32_t)i0, (uint32_t)j0, ok ? J : kNoTile, 0u);
    return;
  }
  // two segments: the last cell lies in the segment after the first cell's
  // (the next plane of the band, or the first plane of the next band), so the
  // block is the end of the first segment, rows j0 .. its band's last row,
  // and the start of the second, rows jB .. j1
  const bool next_band = i0 + 1 == gm.nx;
  const int iB = next_band ? 0 : i0 + 1;
  const uint32_t bandB = next_band ? bandA + 1u : bandA;
  bool ok = i1 == iB && bandL == bandB && gm.nx <= 0xFFFF && gm.ny <= 0xFFFF;
  uint32_t JA = 0u, JB = 0u, jB = 0u;
  if (ok) {
    const uint32_t endA = min((bandA + 1u) * gm.band, (uint32_t)gm.ny) - 1u;
    jB = bandB * gm.band;
    JA = endA - (uint32_t)j0;
    JB = (uint32_t)j1 - jB;
    ok = 3u * (JA + 3u + JB + 3u) * gm.W <= (uint32_t)kWinWords && (JA + 1u + JB + 1u) * gm.W <= 64u;
  }
  tiles[t] = ok ? make_uint4((uint32_t)i0, (uint32_t)j0, JA | (JB << 8) | kTwoSeg, (uint32_t)iB | (jB << 16))
                : make_uint4((uint32_t)i0, (uint32_t)j0, kNoTile, 0u);
}

// face_bits: 16 cells per word, 2 bits per cell
__global__ void __launch_bounds__(kBlock) k_build_face_bits(const KP kp, const int8_t *__restrict__ solid,
                                                            const int32_t *__restrict__ cells,
                                                            uint32_t *__restrict__ face_bits) {
  const uint32_t wi = blockIdx.x * kBlock + threadIdx.x;
  const lbm3d_params &p = kp.p;
  const Geo &gm = kp.geo;
  const uint32_t N = (uint32_t)p.n_cells;
  if (wi >= (N + 15) / 16) return;
  uint32_t word = 0;
  for (uint32_t b = 0; b < 16; ++b) {
    const uint32_t x = wi * 16 + b;
    if (x >= N) break;
    int i, j, k;
    coords(gm, (uint32_t)cells[x], i, j, k);
    const int owner = face_owner(p, gm, i, j, k);
    if (owner < 0) continue;
    int ii = i, jj = j, kk = k;
    const int n = owner < 2 ? gm.nx : (owner < 4 ? gm.ny : gm.nz);
    const int inw = (owner & 1) ? n - 2 : 1;
    if (owner < 2) ii = inw; else if (owner < 4) jj = inw; else kk = inw;
    const bool inward_solid = solid[((int64_t)ii * gm.ny + jj) * gm.nz + kk] > 0;  // SP:278
    word |= (1u | (inward_solid ? 2u : 0u)) << (2 * b);
  }
  face_bits[wi] = word;
}

__global__ void __launch_bounds__(kBlock) k_neighbors(const KP kp, const int32_t *__restrict__ cells,
                                                      const uint4 *__restrict__ rows, int32_t *__restrict__ nbr) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)kp.p.n_cells;
  if (x >= N) return;
  int i, j, k;
  coords(kp.geo, (uint32_t)cells[x], i, j, k);
  int32_t nb[18];
  const Rows rb(rows, kp.geo.n_row_words);
  neighbor_ids(TableRows(rb, kp.geo, i, j), kp.geo, k, nb);
#pragma unroll
  for (int q = 0; q < 18; ++q) nbr[(size_t)q * N + x] = nb[q];
}

// lbm3d_round16: the step kernels' 16-bit store rounding applied to n values
template <typename T>
__global__ void __launch_bounds__(kBlock) k_round16(int stochastic, const float *__restrict__ in,
                                                    const uint32_t *__restrict__ rnd, uint16_t *__restrict__ out,
                                                    uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float v = in[i];
  if constexpr (std::is_same<T, __half>::value) {
    out[i] = stochastic ? half_sr(v, rnd[i] & 0xFFFFu) : __half_as_ushort(__float2half_rn(v));
  } else {
    if (stochastic) {
      out[i] = bf16_sr(v, rnd[i] & 0xFFFFu);
    } else {
      const __hip_bfloat16 h = __float2bfloat16(v);
      out[i] = *reinterpret_cast<const unsigned short *>(&h);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_scatter(const KP kp, const int32_t *__restrict__ cells,
                                                    const float *__restrict__ rho, const float *__restrict__ v,
                                                    int64_t v_stride, int32_t ox, int32_t oy, int32_t oz,
                                                    float *__restrict__ rho_d, float *__restrict__ v_d) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t N = (uint32_t)kp.p.n_cells;
  if (x >= N) return;
  int i, j, k;
  coords(kp.geo, (uint32_t)cells[x], i, j, k);
  const int64_t o = ((int64_t)i * oy + j) * oz + k;
  if (rho_d) rho_d[o] = rho[x];
  if (v_d) {
    v_d[3 * o] = v[x];
    v_d[3 * o + 1] = v[v_stride + x];
    v_d[3 * o + 2] = v[2 * v_stride + x];
  }
}

// FP64 reductions: per-block partials, then one wave per quantity folds them
// in a fixed order (the result does not depend on scheduling).
constexpr int kRed = 8;       // quantities per batch sample
constexpr int kRedGrid = 10;  // quantities of a single-grid reduction

__device__ __forceinline__ double fold(int r, double a, double b) {
  return r == 4 ? (b > a ? b : a) : a + b;  // quantity 4 is a maximum, the others sums
}

// sums over a block: acc[r] of every thread -> out[r] (valid in thread 0..N-1)
template <int N>
__device__ __forceinline__ void block_fold(double acc[N], double *__restrict__ out) {
  __shared__ double s[N][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    double val = acc[r];
    for (int d = 32; d > 0; d >>= 1) val = fold(r, val, __shfl_down(val, d, 64));
    if (lane == 0) s[r][wid] = val;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int r = threadIdx.x;
    double val = s[r][0];
    for (int w = 1; w < kBlock / 64; ++w) val = fold(r, val, s[r][w]);
    out[r] = val;
  }
}

// quantities 0-5: sum rho, sum v (3), max |v|, non-finite cells
template <typename R>
__device__ __forceinline__ bool accumulate(double *acc, R r, R a, R b, R c) {
  const bool fin = isfinite(r) && isfinite(a) && isfinite(b) && isfinite(c);
  if (!fin) {
    acc[5] += 1.0;
    return false;
  }
  acc[0] += r;
  acc[1] += a;
  acc[2] += b;
  acc[3] += c;
  const double nrm = sqrt((double)a * a + (double)b * b + (double)c * c);
  acc[4] = nrm > acc[4] ? nrm : acc[4];
  return true;
}

// 6-8: sum vx, cells and sum rho over the slabs i_lo <= i <= i_hi; 9: 0
// (FP32 fields, or the FP64 fields of FP64 storage)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_reduce_partial(const KP kp, const int32_t *__restrict__ cells,
                                                           const R *__restrict__ rho, const R *__restrict__ v,
                                                           int32_t i_lo, int32_t i_hi, double *__restrict__ part) {
  const uint32_t N = (uint32_t)kp.p.n_cells;
  double acc[kRedGrid] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t x = blockIdx.x * kBlock + threadIdx.x; x < N; x += gridDim.x * kBlock) {
    const R a = v[x], r = rho[x];
    if (!accumulate(acc, r, a, v[(size_t)N + x], v[2 * (size_t)N + x])) continue;
    const int i = (int)fdiv((uint32_t)cells[x], kp.geo.m_nyz, kp.geo.s_nyz);
    if (i >= i_lo && i <= i_hi) {
      acc[6] += a;
      acc[7] += 1.0;
      acc[8] += r;
    }
  }
  double out[kRedGrid];
  block_fold<kRedGrid>(acc, out);
  if (threadIdx.x < kRedGrid) part[(size_t)blockIdx.x * kRedGrid + threadIdx.x] = out[threadIdx.x];
}

// one wave per quantity r: lanes stride over the partials, then a shuffle tree
__global__ void __launch_bounds__(64 * kRedGrid) k_reduce_final(const double *__restrict__ part, int nparts,
                                                               double *__restrict__ out) {
  const int r = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double val = 0.0;
  for (int b = lane; b < nparts; b += 64) val = fold(r, val, part[(size_t)b * kRedGrid + r]);
  for (int d = 32; d > 0; d >>= 1) val = fold(r, val, __shfl_down(val, d, 64));
  if (lane == 0) out[r] = val;
}

// Per-sample reductions of a batch: one block per chunk, then one wave per
// (sample, quantity) over the sample's chunks.  out[s][r]: sum rho, sum v (3),
// max |v|, non-finite cells, fluid cells, 0.
__global__ void __launch_bounds__(kBlock) k_reduce_batch_partial(const Batch bt, const float *__restrict__ rho,
                                                                 const float *__restrict__ v, uint32_t N,
                                                                 double *__restrict__ part) {
  const uint32_t c = blockIdx.x;
  const SampleDesc &d = bt.desc[bt.chunk_sample[c]];
  double acc[kRed] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t x = c * kChunk + threadIdx.x; x < (c + 1) * kChunk && x < d.cell_end; x += kBlock) {
    accumulate(acc, rho[x], v[x], v[(size_t)N + x], v[2 * (size_t)N + x]);
    acc[6] += 1.0;
  }
  double out[kRed];
  block_fold<kRed>(acc, out);
  if (threadIdx.x < kRed) part[(size_t)c * kRed + threadIdx.x] = out[threadIdx.x];
}

__global__ void __launch_bounds__(64 * kRed) k_reduce_batch_final(const Batch bt, const double *__restrict__ part,
                                                                 double *__restrict__ out) {
  const int r = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const SampleDesc &d = bt.desc[blockIdx.x];
  double val = 0.0;
  for (uint32_t c = d.chunk_begin + lane; c < d.chunk_end; c += 64) val = fold(r, val, part[(size_t)c * kRed + r]);
  for (int dd = 32; dd > 0; dd >>= 1) val = fold(r, val, __shfl_down(val, dd, 64));
  if (lane == 0) out[(size_t)blockIdx.x * kRed + r] = val;
}

constexpr int kRedBlocks = 1024;

// ---------------------------------------------------------------------------
// Host-side checks and launches.
// ---------------------------------------------------------------------------
int check_params(const lbm3d_params *p) {
  if (!p) return fail(LBM3D_EINVAL, "%s", "params is NULL");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(LBM3D_EINVAL, "%s", "grid extents must be >= 1");
  if ((int64_t)p->nx * p->ny * p->nz > INT32_MAX)
    return fail(LBM3D_ESIZE, "%s", "grid has more than 2^31-1 cells");
  if (p->band_rows < 0) return fail(LBM3D_EINVAL, "%s", "band_rows must be >= 0");
  for (int f = 0; f < 6; ++f) {
    if (p->bc_mode[f] == 1 || p->bc_mode[f] == 2) {
      const int n = f < 2 ? p->nx : (f < 4 ? p->ny : p->nz);
      if (n < 2) return fail(LBM3D_EINVAL, "%s", "a face with a boundary condition needs extent >= 2");
    }
  }
  return LBM3D_OK;
}

int check_cells(const lbm3d_params *p) {
  if (p->n_cells < 0 || (int64_t)p->n_cells > (int64_t)p->nx * p->ny * p->nz)
    return fail(LBM3D_EINVAL, "%s", "n_cells out of range");
  // a pair of FP32 slots is one buffer resource: its bytes must fit the 32-bit record count
  if ((uint64_t)slot_stride((uint32_t)p->n_cells) * 2 * sizeof(float) > UINT32_MAX)
    return fail(LBM3D_ESIZE, "%s", "two slots of n_cells exceed a 32-bit buffer range");
  return LBM3D_OK;
}

int check_dtype_size(const lbm3d_params *p, int dtype) {
  const uint64_t elt = dtype == LBM3D_F64 ? 8 : (dtype == LBM3D_F32 ? 4 : 2);
  if ((uint64_t)slot_stride((uint32_t)p->n_cells) * 2 * elt > UINT32_MAX)
    return fail(LBM3D_ESIZE, "%s", "two slots of n_cells exceed a 32-bit buffer range for this dtype");
  return LBM3D_OK;
}

int check_face_offsets(const lbm3d_params *p) {
  for (int f = 0; f < 6; ++f)
    if (p->bc_mode[f] == 1 && p->face_u_off[f] < 0)
      return fail(LBM3D_EINVAL, "%s", "a pressure face has no face_u_off");
  return LBM3D_OK;
}

bool valid_dtype(int d) { return d == LBM3D_F32 || d == LBM3D_F16 || d == LBM3D_BF16 || d == LBM3D_F64; }

int check_dtype_size(const lbm3d_params *p, int dtype);

int solver_args(const lbm3d_params *p, int dtype, const void *g, const int32_t *cells, const void *rows,
                const uint32_t *face_bits) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if ((rc = check_face_offsets(p))) return rc;
  if (!valid_dtype(dtype)) return fail(LBM3D_EINVAL, "%s", "unknown dtype");
  if ((rc = check_dtype_size(p, dtype))) return rc;
  if (p->n_cells > 0 && (!g || !cells || !rows || !face_bits))
    return fail(LBM3D_EINVAL, "%s", "populations, cells, rows or face_bits is NULL");
  return LBM3D_OK;
}

// seed of step t's stochastic rounding (0: round to nearest)
uint32_t sr_seed(const lbm3d_params &p, int64_t t) {
  if (!(p.flags & LBM3D_FLAG_STOCHASTIC_ROUNDING)) return 0u;
  uint64_t h = (uint64_t)t * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  h ^= h >> 31;
  h *= 0xD6E8FEB86659FD93ull;
  h ^= h >> 32;
  return (uint32_t)h | 1u;
}

// step t over blocks lb_begin .. lb_begin + nb - 1
template <typename T, bool FORCED>
void launch_step_parity(hipStream_t s, KP kp, void *g, const int32_t *cells, const uint4 *rows,
                        const uint32_t *face_bits, void *face_u, int64_t t, uint32_t lb_begin, uint32_t nb) {
  kp.sr_seed = sr_seed(kp.p, t);
  using R = typename St<T>::R;
  const unsigned grid = (unsigned)xcd_grid(nb, (t & 1) ? LBM3D_XCD_CHUNK_ODD : LBM3D_XCD_CHUNK_EVEN);
  if (t & 1)
    k_step<T, true, FORCED><<<grid, kBlock, 0, s>>>(kp, (T *)g, cells, rows, face_bits, (R *)face_u, t == 0,
                                                    lb_begin, nb);
  else
    k_step<T, false, FORCED><<<grid, kBlock, 0, s>>>(kp, (T *)g, cells, rows, face_bits, (R *)face_u, t == 0,
                                                     lb_begin, nb);
}

uint32_t step_blocks(const lbm3d_params &p) { return ((uint32_t)p.n_cells + kBlock - 1) / kBlock; }

// step t over the blocks lb_begin .. lb_begin + nb - 1 (nb = 0: the whole list)
template <typename T>
int launch_step(hipStream_t s, const KP &kp, void *g, const int32_t *cells, const uint4 *rows,
                const uint32_t *face_bits, void *face_u, int64_t t, uint32_t lb_begin = 0, uint32_t nb = 0) {
  if (kp.p.n_cells == 0) return LBM3D_OK;
  if (nb == 0) nb = step_blocks(kp.p);
  if (kp.p.force_on)
    launch_step_parity<T, true>(s, kp, g, cells, rows, face_bits, face_u, t, lb_begin, nb);
  else
    launch_step_parity<T, false>(s, kp, g, cells, rows, face_bits, face_u, t, lb_begin, nb);
  return check_launch("lbm3d_step");
}

template <typename T>
int launch_macros(hipStream_t s, const KP &kp, const void *g, const int32_t *cells, const uint4 *rows,
                  const uint32_t *face_bits, const void *face_u, int64_t t, void *rho, void *v) {
  using R = typename St<T>::R;
  const unsigned grid = (unsigned)((kp.p.n_cells + kBlock - 1) / kBlock);
  if (grid == 0) return LBM3D_OK;
  if (t & 1)
    k_macros<T, true><<<grid, kBlock, 0, s>>>(kp, (const T *)g, cells, rows, face_bits, (const R *)face_u, t == 0,
                                              (R *)rho, (R *)v);
  else
    k_macros<T, false><<<grid, kBlock, 0, s>>>(kp, (const T *)g, cells, rows, face_bits, (const R *)face_u, t == 0,
                                               (R *)rho, (R *)v);
  return check_launch("lbm3d_macros");
}

template <typename T>
int launch_init(hipStream_t s, const KP &kp, void *g, const int32_t *cells, void *face_u, const float *rho0,
                const float *u0) {
  using R = typename St<T>::R;
  const unsigned grid = (unsigned)((kp.p.n_cells + kBlock - 1) / kBlock);
  if (grid == 0) return LBM3D_OK;
  k_init<T><<<grid, kBlock, 0, s>>>(kp, (T *)g, cells, (R *)face_u, rho0, u0);
  return check_launch("lbm3d_init");
}

// ---- batches ----------------------------------------------------------------
constexpr size_t kDescAlign = 256;

size_t desc_table_bytes(int32_t n) {
  return ((size_t)n * sizeof(SampleDesc) + kDescAlign - 1) / kDescAlign * kDescAlign;
}

Batch batch_view(const void *desc, int32_t n_samples, const uint8_t *active) {
  Batch b;
  b.desc = (const SampleDesc *)desc;
  b.chunk_sample = (const int32_t *)((const char *)desc + desc_table_bytes(n_samples));
  b.active = active;
  return b;
}

// compute units of the current device, cached per device
int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
  if (dev < 64) {
    const int c = cache[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  if (dev < 64) cache[dev].store(cus, std::memory_order_relaxed);
  return cus;
}

constexpr lbm3d_launch kDefaultLaunch[2] = {{kBlock, 0, LBM3D_REVERSE_EVEN, 0}, {kBlock, 0, 0, 0}};

bool valid_launch(const lbm3d_launch &c) {
  return (c.block == 128 || c.block == 256 || c.block == 512) && c.grid_per_cu >= 0 && c.grid_per_cu <= 64;
}

template <typename T, bool ODD, bool FORCED>
void launch_batch_parity(hipStream_t s, const KP &kp, const Batch &bt, void *g, const int32_t *cells,
                         const uint4 *rows, const uint32_t *face_bits, void *face_u, int first,
                         const lbm3d_launch &c, uint32_t cell_begin, uint32_t cell_end) {
  using R = typename St<T>::R;
  // cell_begin and cell_end are multiples of kChunk, so of every block size
  const uint32_t tile0 = cell_begin / (uint32_t)c.block;
  const uint32_t n_tiles = (cell_end - cell_begin) / (uint32_t)c.block;
  constexpr uint32_t C = ODD ? LBM3D_BATCH_XCD_CHUNK_ODD : LBM3D_BATCH_XCD_CHUNK_EVEN;
  // XCD runs need a grid of 8 * per_xcd workgroups
  uint32_t grid = C > 0 ? (uint32_t)xcd_grid(n_tiles, C) : n_tiles;
  if (c.grid_per_cu > 0) {
    const uint64_t g_max = (uint64_t)device_cus() * (uint64_t)c.grid_per_cu;
    if (g_max < grid) grid = C > 0 ? (uint32_t)std::max<uint64_t>(8, g_max / 8 * 8) : (uint32_t)g_max;
  }
  const int rev = c.reverse ? 1 : 0;
#define LBM3D_BATCH_LAUNCH(B)                                                                             \
  if (c.grid_per_cu > 0)                                                                                  \
    k_step_batch<T, ODD, FORCED, B><<<grid, B, 0, s>>>(kp, bt, (T *)g, cells, rows, face_bits, (R *)face_u,  \
                                                       first, tile0, n_tiles, rev);                        \
  else                                                                                                    \
    k_step_batch1<T, ODD, FORCED, B><<<grid, B, 0, s>>>(kp, bt, (T *)g, cells, rows, face_bits, (R *)face_u, \
                                                        first, tile0, n_tiles, rev)
  switch (c.block) {
    case 128: LBM3D_BATCH_LAUNCH(128); break;
    case 512: LBM3D_BATCH_LAUNCH(512); break;
    default: LBM3D_BATCH_LAUNCH(256); break;
  }
#undef LBM3D_BATCH_LAUNCH
}

template <typename T>
int launch_batch_step(hipStream_t s, const KP &kp, const Batch &bt, void *g, const int32_t *cells, const uint4 *rows,
                      const uint32_t *face_bits, void *face_u, int64_t t, const lbm3d_launch *cfg, uint32_t cb,
                      uint32_t ce) {
  const lbm3d_launch &c = cfg ? cfg[t & 1] : kDefaultLaunch[t & 1];
  const int first = t == 0;
  KP kpt = kp;
  kpt.sr_seed = sr_seed(kp.p, t);
  if (kp.p.force_on) {
    if (t & 1) launch_batch_parity<T, true, true>(s, kpt, bt, g, cells, rows, face_bits, face_u, first, c, cb, ce);
    else launch_batch_parity<T, false, true>(s, kpt, bt, g, cells, rows, face_bits, face_u, first, c, cb, ce);
  } else {
    if (t & 1) launch_batch_parity<T, true, false>(s, kpt, bt, g, cells, rows, face_bits, face_u, first, c, cb, ce);
    else launch_batch_parity<T, false, false>(s, kpt, bt, g, cells, rows, face_bits, face_u, first, c, cb, ce);
  }
  return check_launch("lbm3d_batch_step");
}

template <typename T>
int launch_batch_macros(hipStream_t s, const KP &kp, const Batch &bt, const void *g, const int32_t *cells,
                        const uint4 *rows, const uint32_t *face_bits, const void *face_u, int64_t t, void *rho,
                        void *v) {
  using R = typename St<T>::R;
  const unsigned grid = (unsigned)((uint32_t)kp.p.n_cells / kBlock);
  if (t & 1)
    k_macros_batch<T, true><<<grid, kBlock, 0, s>>>(kp, bt, (const T *)g, cells, rows, face_bits, (const R *)face_u,
                                                    t == 0, (R *)rho, (R *)v);
  else
    k_macros_batch<T, false><<<grid, kBlock, 0, s>>>(kp, bt, (const T *)g, cells, rows, face_bits, (const R *)face_u,
                                                     t == 0, (R *)rho, (R *)v);
  return check_launch("lbm3d_batch_macros");
}

template <typename T>
int launch_batch_init(hipStream_t s, const KP &kp, const Batch &bt, void *g, const int32_t *cells, void *face_u,
                      const float *rho0, const float *u0) {
  using R = typename St<T>::R;
  const unsigned grid = (unsigned)((uint32_t)kp.p.n_cells / kBlock);
  k_init_batch<T><<<grid, kBlock, 0, s>>>(kp, bt, (T *)g, cells, (R *)face_u, rho0, u0);
  return check_launch("lbm3d_batch_init");
}

// physical parameters and the batch-wide cell count of a batch call
int check_batch(const lbm3d_params *p, int32_t n_samples, const void *desc) {
  if (!p) return fail(LBM3D_EINVAL, "%s", "params is NULL");
  if (n_samples < 1 || !desc) return fail(LBM3D_EINVAL, "%s", "a batch needs n_samples >= 1 and a descriptor");
  if (p->n_cells <= 0 || p->n_cells % (int32_t)kChunk != 0)
    return fail(LBM3D_EINVAL, "%s", "batch n_cells must be the layout's cell count (a positive multiple of "
                                    "LBM3D_BATCH_ALIGN)");
  if (p->n_cells / (int32_t)kChunk < n_samples)
    return fail(LBM3D_EINVAL, "%s", "batch n_cells is smaller than one chunk per sample");
  return LBM3D_OK;
}

int check_batch_solver(const lbm3d_params *p, int dtype, int32_t n_samples, const void *desc, const void *g,
                       const int32_t *cells, const void *rows, const uint32_t *face_bits) {
  int rc = check_batch(p, n_samples, desc);
  if (rc) return rc;
  if (!valid_dtype(dtype)) return fail(LBM3D_EINVAL, "%s", "unknown dtype");
  if ((rc = check_dtype_size(p, dtype))) return rc;
  if (!g || !cells || !rows || !face_bits)
    return fail(LBM3D_EINVAL, "%s", "populations, cells, rows or face_bits is NULL");
  return LBM3D_OK;
}

KP make_kp_batch(const lbm3d_params &p) {
  KP k;
  k.p = p;
  fill_phys(k.pf, p);
  fill_phys(k.pd, p);
  k.S = slot_stride((uint32_t)p.n_cells);
  memset(&k.geo, 0, sizeof(k.geo));  // every sample's geometry comes from its descriptor
  k.sr_seed = 0u;
  return k;
}

// Scratch of lbm3d_reduce_noalloc: the per-block partials of the reduction
// (kRedBlocks x kRedGrid doubles, 80 KiB), a file-scope device array, so the
// entry point allocates nothing.  Each GPU holds its own copy at its own
// address; host code must use the address hipGetSymbolAddress returns on that
// device, never the array's name.  Calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[kRedBlocks * kRedGrid];

constexpr int kMaxDevices = 64;
std::atomic<double *> g_scratch_ptr[kMaxDevices];

// The current device's address of `scratch`, looked up once per device.
int reduce_scratch(double **ptr) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(LBM3D_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(LBM3D_EINVAL, "%s", "device index beyond the scratch table");
  double *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    void *addr = nullptr;
    e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !addr) return fail(LBM3D_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    p = (double *)addr;
    g_scratch_ptr[dev].store(p, std::memory_order_release);
  }
  *ptr = p;
  return LBM3D_OK;
}

size_t reduce_workspace_bytes() { return (size_t)kRedBlocks * kRedGrid * sizeof(double); }

// The two passes of lbm3d_reduce / lbm3d_reduce64 over fields of type R.
template <typename R>
int launch_reduce(void *stream, const lbm3d_params *p, const int32_t *cells, const R *rho, const R *v, int32_t i_lo,
                  int32_t i_hi, double *out, void *workspace, size_t workspace_bytes, const char *what) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!out || !workspace || (p->n_cells > 0 && (!cells || !rho || !v)))
    return fail(LBM3D_EINVAL, "NULL argument to %s", what);
  if (workspace_bytes < reduce_workspace_bytes()) return fail(LBM3D_EINVAL, "%s: workspace too small", what);
  int blocks = (p->n_cells + kBlock - 1) / kBlock;
  blocks = blocks < 1 ? 1 : (blocks > kRedBlocks ? kRedBlocks : blocks);
  const KP kp = make_kp(*p);
  hipStream_t s = (hipStream_t)stream;
  k_reduce_partial<R><<<blocks, kBlock, 0, s>>>(kp, cells, rho, v, i_lo, i_hi, (double *)workspace);
  if ((rc = check_launch(what))) return rc;
  k_reduce_final<<<1, 64 * kRedGrid, 0, s>>>((const double *)workspace, blocks, out);
  return check_launch(what);
}

}  // namespace

extern "C" {

int lbm3d_abi_version(void) { return LBM3D_ABI_VERSION; }

const char *lbm3d_last_error(void) { return g_err; }

int64_t lbm3d_slot_stride(int32_t n_cells) { return n_cells < 0 ? 0 : (int64_t)slot_stride((uint32_t)n_cells); }

int64_t lbm3d_row_words(const lbm3d_params *p) {
  if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) return 0;
  return (int64_t)p->nx * p->ny * ((p->nz + 31) / 32);
}

int64_t lbm3d_face_bit_words(int32_t n_cells) { return n_cells < 0 ? 0 : ((int64_t)n_cells + 15) / 16; }

int64_t lbm3d_face_plane(const lbm3d_params *p, int32_t face) {
  if (!p || face < 0 || face > 5) return 0;
  return face < 2 ? (int64_t)p->ny * p->nz : (face < 4 ? (int64_t)p->nx * p->nz : (int64_t)p->nx * p->ny);
}

size_t lbm3d_compact_workspace_bytes(int64_t n_dense) {
  const int64_t tiles = (n_dense + kTile - 1) / kTile;
  return (size_t)(tiles > 0 ? tiles : 1) * sizeof(int32_t);
}

int lbm3d_compact_count(void *stream, const lbm3d_params *p, const int8_t *solid, int32_t *count, void *workspace,
                        size_t workspace_bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!solid || !count || !workspace) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_compact_count");
  const int64_t n = (int64_t)p->nx * p->ny * p->nz;
  if (workspace_bytes < lbm3d_compact_workspace_bytes(n))
    return fail(LBM3D_EINVAL, "%s", "compaction workspace too small");
  const int64_t tiles = (n + kTile - 1) / kTile;
  hipStream_t s = (hipStream_t)stream;
  k_compact_count<<<(unsigned)tiles, kBlock, 0, s>>>(solid, list_order(*p), n, (int32_t *)workspace);
  if ((rc = check_launch("lbm3d_compact_count"))) return rc;
  k_compact_scan<<<1, kBlock, 0, s>>>((int32_t *)workspace, tiles, count);
  return check_launch("lbm3d_compact_count/scan");
}

int lbm3d_compact_fill(void *stream, const lbm3d_params *p, const int8_t *solid, int32_t *d2c, int32_t *cells,
                       void *workspace, size_t workspace_bytes) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!solid || !workspace || (p->n_cells > 0 && !cells))
    return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_compact_fill");
  const int64_t n = (int64_t)p->nx * p->ny * p->nz;
  if (workspace_bytes < lbm3d_compact_workspace_bytes(n))
    return fail(LBM3D_EINVAL, "%s", "compaction workspace too small");
  const int64_t tiles = (n + kTile - 1) / kTile;
  k_compact_fill<<<(unsigned)tiles, kBlock, 0, (hipStream_t)stream>>>(solid, list_order(*p), n,
                                                                      (const int32_t *)workspace, d2c, cells);
  return check_launch("lbm3d_compact_fill");
}

int lbm3d_build_rows_at(void *stream, const lbm3d_params *p, const int8_t *solid, const int32_t *d2c,
                        int32_t id_offset, lbm3d_row_word *rows) {
  int rc = check_params(p);
  if (rc) return rc;
  if (!solid || !d2c || !rows) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_build_rows");
  if (id_offset < 0) return fail(LBM3D_EINVAL, "%s", "negative id offset");
  if ((int64_t)id_offset + (int64_t)p->nx * p->ny * p->nz > INT32_MAX)
    return fail(LBM3D_ESIZE, "%s", "id offset plus cells exceeds int32");
  const Geo gm = make_geo(p->nx, p->ny, p->nz, p->face_u_off, p->band_rows);
  const int64_t n = lbm3d_row_words(p);
  k_build_rows<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, (hipStream_t)stream>>>(gm, solid, d2c, id_offset,
                                                                                         (uint4 *)rows);
  return check_launch("lbm3d_build_rows");
}

int lbm3d_build_rows(void *stream, const lbm3d_params *p, const int8_t *solid, const int32_t *d2c,
                     lbm3d_row_word *rows) {
  return lbm3d_build_rows_at(stream, p, solid, d2c, 0, rows);
}

int64_t lbm3d_tile_words(const lbm3d_params *p) {
  if (!p || p->n_cells <= 0) return 0;
  return ((int64_t)p->n_cells + kBlock - 1) / kBlock;
}

int lbm3d_build_tiles(void *stream, const lbm3d_params *p, const int32_t *cells, lbm3d_row_word *tiles) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (p->n_cells == 0) return LBM3D_OK;
  if (!cells || !tiles) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_build_tiles");
  const Geo gm = make_geo(p->nx, p->ny, p->nz, p->face_u_off, p->band_rows);
  const int64_t nt = lbm3d_tile_words(p);
  k_build_tiles<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, (hipStream_t)stream>>>(
      gm, cells, (uint32_t)p->n_cells, (uint4 *)tiles);
  return check_launch("lbm3d_build_tiles");
}

int lbm3d_build_face_bits(void *stream, const lbm3d_params *p, const int8_t *solid, const int32_t *cells,
                          uint32_t *face_bits) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (p->n_cells == 0) return LBM3D_OK;
  if (!solid || !cells || !face_bits) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_build_face_bits");
  const KP kp = make_kp(*p);
  const int64_t words = lbm3d_face_bit_words(p->n_cells);
  k_build_face_bits<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, (hipStream_t)stream>>>(kp, solid, cells,
                                                                                                  face_bits);
  return check_launch("lbm3d_build_face_bits");
}

int lbm3d_build_neighbors(void *stream, const lbm3d_params *p, const int32_t *cells, const lbm3d_row_word *rows,
                          int32_t *nbr) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (p->n_cells == 0) return LBM3D_OK;
  if (!cells || !rows || !nbr) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_build_neighbors");
  const KP kp = make_kp(*p);
  k_neighbors<<<(unsigned)((p->n_cells + kBlock - 1) / kBlock), kBlock, 0, (hipStream_t)stream>>>(
      kp, cells, (const uint4 *)rows, nbr);
  return check_launch("lbm3d_build_neighbors");
}

int lbm3d_init(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *cells, void *face_u,
               const float *rho0, const float *u0) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if ((rc = check_face_offsets(p))) return rc;
  if (!valid_dtype(dtype)) return fail(LBM3D_EINVAL, "%s", "unknown dtype");
  if ((rc = check_dtype_size(p, dtype))) return rc;
  if (p->n_cells > 0 && (!g || !cells)) return fail(LBM3D_EINVAL, "%s", "populations or cells are NULL");
  for (int f = 0; f < 6; ++f)
    if (p->bc_mode[f] == 1 && !face_u) return fail(LBM3D_EINVAL, "%s", "face_u is NULL with a pressure face");
  const KP kp = make_kp(*p);
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case LBM3D_F32: return launch_init<float>(s, kp, g, cells, face_u, rho0, u0);
    case LBM3D_F16: return launch_init<__half>(s, kp, g, cells, face_u, rho0, u0);
    case LBM3D_F64: return launch_init<double>(s, kp, g, cells, face_u, rho0, u0);
    default: return launch_init<__hip_bfloat16>(s, kp, g, cells, face_u, rho0, u0);
  }
}

int lbm3d_step(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *cells,
               const lbm3d_row_word *rows, const uint32_t *face_bits, void *face_u, int64_t t) {
  int rc = solver_args(p, dtype, g, cells, rows, face_bits);
  if (rc) return rc;
  if (t < 0) return fail(LBM3D_EINVAL, "%s", "negative step index");
  const KP kp = make_kp(*p);
  hipStream_t s = (hipStream_t)stream;
  const uint4 *r = (const uint4 *)rows;
  switch (dtype) {
    case LBM3D_F32: return launch_step<float>(s, kp, g, cells, r, face_bits, face_u, t);
    case LBM3D_F16: return launch_step<__half>(s, kp, g, cells, r, face_bits, face_u, t);
    case LBM3D_F64: return launch_step<double>(s, kp, g, cells, r, face_bits, face_u, t);
    default: return launch_step<__hip_bfloat16>(s, kp, g, cells, r, face_bits, face_u, t);
  }
}

int lbm3d_run(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *cells,
              const lbm3d_row_word *rows, const uint32_t *face_bits, void *face_u, int64_t t0, int32_t nsteps) {
  int rc = solver_args(p, dtype, g, cells, rows, face_bits);
  if (rc) return rc;
  if (t0 < 0 || nsteps < 0) return fail(LBM3D_EINVAL, "%s", "negative step index or count");
  const KP kp = make_kp(*p);
  hipStream_t s = (hipStream_t)stream;
  const uint4 *r = (const uint4 *)rows;
  for (int32_t k = 0; k < nsteps; ++k) {
    switch (dtype) {
      case LBM3D_F32: rc = launch_step<float>(s, kp, g, cells, r, face_bits, face_u, t0 + k); break;
      case LBM3D_F16: rc = launch_step<__half>(s, kp, g, cells, r, face_bits, face_u, t0 + k); break;
      case LBM3D_F64: rc = launch_step<double>(s, kp, g, cells, r, face_bits, face_u, t0 + k); break;
      default: rc = launch_step<__hip_bfloat16>(s, kp, g, cells, r, face_bits, face_u, t0 + k); break;
    }
    if (rc) return rc;
  }
  return LBM3D_OK;
}

int lbm3d_macros(void *stream, const lbm3d_params *p, int dtype, const void *g, const int32_t *cells,
                 const lbm3d_row_word *rows, const uint32_t *face_bits, const void *face_u, int64_t t, void *rho,
                 void *v) {
  int rc = solver_args(p, dtype, g, cells, rows, face_bits);
  if (rc) return rc;
  if (t < 0) return fail(LBM3D_EINVAL, "%s", "negative step index");
  if (p->n_cells > 0 && (!rho || !v)) return fail(LBM3D_EINVAL, "%s", "rho or v output is NULL");
  const KP kp = make_kp(*p);
  hipStream_t s = (hipStream_t)stream;
  const uint4 *r = (const uint4 *)rows;
  switch (dtype) {
    case LBM3D_F32: return launch_macros<float>(s, kp, g, cells, r, face_bits, face_u, t, rho, v);
    case LBM3D_F16: return launch_macros<__half>(s, kp, g, cells, r, face_bits, face_u, t, rho, v);
    case LBM3D_F64: return launch_macros<double>(s, kp, g, cells, r, face_bits, face_u, t, rho, v);
    default: return launch_macros<__hip_bfloat16>(s, kp, g, cells, r, face_bits, face_u, t, rho, v);
  }
}

int lbm3d_scatter_dense(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho,
                        const float *v, int64_t v_stride, int32_t ox, int32_t oy, int32_t oz, float *rho_dense,
                        float *v_dense) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (ox < p->nx || oy < p->ny || oz < p->nz) return fail(LBM3D_EINVAL, "%s", "dense extent smaller than the grid");
  if (v_stride == 0) v_stride = p->n_cells;
  if (v_stride < p->n_cells) return fail(LBM3D_EINVAL, "%s", "v_stride smaller than n_cells");
  if (p->n_cells == 0) return LBM3D_OK;
  if (!cells || (rho_dense && !rho) || (v_dense && !v))
    return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_scatter_dense");
  const KP kp = make_kp(*p);
  const unsigned grid = (unsigned)((p->n_cells + kBlock - 1) / kBlock);
  k_scatter<<<grid, kBlock, 0, (hipStream_t)stream>>>(kp, cells, rho, v, v_stride, ox, oy, oz, rho_dense, v_dense);
  return check_launch("lbm3d_scatter_dense");
}

int lbm3d_round16(void *stream, int dtype, int stochastic, const float *in, const uint32_t *rnd, uint16_t *out,
                  int64_t n) {
  if (dtype != LBM3D_F16 && dtype != LBM3D_BF16) return fail(LBM3D_EINVAL, "%s", "lbm3d_round16: dtype must be F16 or BF16");
  if (n < 0 || n > INT32_MAX) return fail(LBM3D_EINVAL, "%s", "lbm3d_round16: n out of range");
  if (n == 0) return LBM3D_OK;
  if (!in || !out || (stochastic && !rnd)) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_round16");
  const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
  if (dtype == LBM3D_F16)
    k_round16<__half><<<grid, kBlock, 0, (hipStream_t)stream>>>(stochastic, in, rnd, out, (uint32_t)n);
  else
    k_round16<__hip_bfloat16><<<grid, kBlock, 0, (hipStream_t)stream>>>(stochastic, in, rnd, out, (uint32_t)n);
  return check_launch("lbm3d_round16");
}

size_t lbm3d_reduce_workspace_bytes(int32_t n_cells) {
  (void)n_cells;
  return reduce_workspace_bytes();
}

int lbm3d_reduce(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho, const float *v,
                 int32_t i_lo, int32_t i_hi, double *out, void *workspace, size_t workspace_bytes) {
  return launch_reduce<float>(stream, p, cells, rho, v, i_lo, i_hi, out, workspace, workspace_bytes, "lbm3d_reduce");
}

int lbm3d_reduce64(void *stream, const lbm3d_params *p, const int32_t *cells, const double *rho, const double *v,
                   int32_t i_lo, int32_t i_hi, double *out, void *workspace, size_t workspace_bytes) {
  return launch_reduce<double>(stream, p, cells, rho, v, i_lo, i_hi, out, workspace, workspace_bytes,
                               "lbm3d_reduce64");
}

int lbm3d_reduce_noalloc(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho, const float *v,
                         int32_t i_lo, int32_t i_hi, double *out) {
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_cells(p))) return rc;
  if (!out || (p->n_cells > 0 && (!cells || !rho || !v)))
    return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_reduce_noalloc");
  double *part = nullptr;
  if ((rc = reduce_scratch(&part))) return rc;
  int blocks = (p->n_cells + kBlock - 1) / kBlock;
  blocks = blocks < 1 ? 1 : (blocks > kRedBlocks ? kRedBlocks : blocks);
  const KP kp = make_kp(*p);
  hipStream_t s = (hipStream_t)stream;
  // every partial the final pass reads is written by this call's first pass
  k_reduce_partial<float><<<blocks, kBlock, 0, s>>>(kp, cells, rho, v, i_lo, i_hi, part);
  if ((rc = check_launch("lbm3d_reduce_noalloc"))) return rc;
  k_reduce_final<<<1, 64 * kRedGrid, 0, s>>>(part, blocks, out);
  return check_launch("lbm3d_reduce_noalloc/final");
}

// ---- batches ---------------------------------------------------------------
int lbm3d_batch_layout(const lbm3d_params *p, lbm3d_sample *samples, int32_t n_samples, lbm3d_batch_sizes *sizes) {
  if (!p || !samples || !sizes || n_samples < 1) return fail(LBM3D_EINVAL, "%s", "NULL argument or no samples");
  int64_t cell = 0, rows = 0, fu = 0;
  for (int32_t s = 0; s < n_samples; ++s) {
    lbm3d_sample &sm = samples[s];
    lbm3d_params q = *p;
    q.nx = sm.nx;
    q.ny = sm.ny;
    q.nz = sm.nz;
    q.n_cells = sm.n_cells;
    int rc = check_params(&q);
    if (rc) return rc;
    if ((rc = check_cells(&q))) return rc;
    sm.cell_begin = (int32_t)cell;
    sm.rows_begin = rows;
    sm.face_u_begin = (int32_t)fu;
    // every sample owns at least one chunk, so the chunk map names each
    const int64_t span = sm.n_cells > 0 ? sm.n_cells : 1;
    cell += (span + kChunk - 1) / kChunk * kChunk;
    rows += lbm3d_row_words(&q);
    for (int f = 0; f < 6; ++f)
      if (p->bc_mode[f] == 1) fu += 3 * lbm3d_face_plane(&q, f);
    if (cell > INT32_MAX - (int64_t)kChunk || fu > INT32_MAX || rows > UINT32_MAX)
      return fail(LBM3D_ESIZE, "%s", "batch exceeds 2^31 cells or face nodes");
  }
  sizes->cells = cell;
  sizes->row_words = rows;
  sizes->face_u = fu;
  sizes->desc_bytes = (int64_t)desc_table_bytes(n_samples) + (cell / kChunk) * (int64_t)sizeof(int32_t);
  return LBM3D_OK;
}

int lbm3d_batch_desc(const lbm3d_params *p, const lbm3d_sample *samples, int32_t n_samples, void *host_desc,
                     size_t bytes) {
  if (!p || !samples || !host_desc || n_samples < 1) return fail(LBM3D_EINVAL, "%s", "NULL argument or no samples");
  const lbm3d_sample &last = samples[n_samples - 1];
  const int64_t span = last.n_cells > 0 ? last.n_cells : 1;
  const int64_t cells = last.cell_begin + (span + kChunk - 1) / kChunk * kChunk;
  const int64_t n_chunks = cells / kChunk;
  if (bytes < desc_table_bytes(n_samples) + (size_t)n_chunks * sizeof(int32_t))
    return fail(LBM3D_EINVAL, "%s", "descriptor buffer too small");
  memset(host_desc, 0, desc_table_bytes(n_samples));
  SampleDesc *d = (SampleDesc *)host_desc;
  int32_t *map = (int32_t *)((char *)host_desc + desc_table_bytes(n_samples));
  int64_t expect = 0, total_rows = 0;
  for (int32_t s = 0; s < n_samples; ++s) {
    const lbm3d_sample &sm = samples[s];
    if (sm.cell_begin != expect) return fail(LBM3D_EINVAL, "%s", "samples not placed by lbm3d_batch_layout");
    int32_t off[6];
    int64_t fu = sm.face_u_begin;
    lbm3d_params q = *p;
    q.nx = sm.nx;
    q.ny = sm.ny;
    q.nz = sm.nz;
    q.n_cells = sm.n_cells;
    int rc = check_params(&q);
    if (rc) return rc;
    if ((rc = check_cells(&q))) return rc;
    for (int f = 0; f < 6; ++f) {
      off[f] = p->bc_mode[f] == 1 ? (int32_t)fu : -1;
      if (p->bc_mode[f] == 1) fu += 3 * lbm3d_face_plane(&q, f);
    }
    d[s].geo = make_geo(sm.nx, sm.ny, sm.nz, off, p->band_rows);
    d[s].cell_begin = (uint32_t)sm.cell_begin;
    d[s].cell_end = (uint32_t)(sm.cell_begin + (sm.n_cells > 0 ? sm.n_cells : 0));
    const int64_t sp = sm.n_cells > 0 ? sm.n_cells : 1;
    const int64_t c0 = sm.cell_begin / kChunk, c1 = c0 + (sp + kChunk - 1) / kChunk;
    d[s].chunk_begin = (uint32_t)c0;
    d[s].chunk_end = (uint32_t)c1;
    d[s].rows_begin = (uint64_t)sm.rows_begin;
    for (int64_t c = c0; c < c1; ++c) map[c] = s;
    expect = c1 * kChunk;
    total_rows = sm.rows_begin + lbm3d_row_words(&q);
  }
  // tile descriptors follow the whole row table, one per kBlock cells of the batch
  for (int32_t s = 0; s < n_samples; ++s) d[s].tiles_begin = (uint64_t)total_rows + d[s].cell_begin / kBlock;
  return LBM3D_OK;
}

int lbm3d_batch_init(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc, int32_t n_samples,
                     const int32_t *cells, void *face_u, const float *rho0, const float *u0) {
  int rc = check_batch(p, n_samples, desc);
  if (rc) return rc;
  if (!valid_dtype(dtype)) return fail(LBM3D_EINVAL, "%s", "unknown dtype");
  if ((rc = check_dtype_size(p, dtype))) return rc;
  if (!g || !cells) return fail(LBM3D_EINVAL, "%s", "populations or cells are NULL");
  for (int f = 0; f < 6; ++f)
    if (p->bc_mode[f] == 1 && !face_u) return fail(LBM3D_EINVAL, "%s", "face_u is NULL with a pressure face");
  const KP kp = make_kp_batch(*p);
  const Batch bt = batch_view(desc, n_samples, nullptr);
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case LBM3D_F32: return launch_batch_init<float>(s, kp, bt, g, cells, face_u, rho0, u0);
    case LBM3D_F16: return launch_batch_init<__half>(s, kp, bt, g, cells, face_u, rho0, u0);
    case LBM3D_F64: return launch_batch_init<double>(s, kp, bt, g, cells, face_u, rho0, u0);
    default: return launch_batch_init<__hip_bfloat16>(s, kp, bt, g, cells, face_u, rho0, u0);
  }
}

int lbm3d_batch_run(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc, int32_t n_samples,
                    const int32_t *cells, const lbm3d_row_word *rows, const uint32_t *face_bits, void *face_u,
                    const uint8_t *active, int64_t t0, int32_t nsteps, const lbm3d_launch *cfg) {
  if (!p) return fail(LBM3D_EINVAL, "%s", "params is NULL");
  return lbm3d_batch_run_range(stream, p, dtype, g, desc, n_samples, cells, rows, face_bits, face_u, active, t0,
                               nsteps, 0, p->n_cells, cfg);
}

int lbm3d_batch_run_range(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc,
                          int32_t n_samples, const int32_t *cells, const lbm3d_row_word *rows,
                          const uint32_t *face_bits, void *face_u, const uint8_t *active, int64_t t0, int32_t nsteps,
                          int32_t cell_begin, int32_t cell_end, const lbm3d_launch *cfg) {
  int rc = check_batch_solver(p, dtype, n_samples, desc, g, cells, rows, face_bits);
  if (rc) return rc;
  if (t0 < 0 || nsteps < 0) return fail(LBM3D_EINVAL, "%s", "negative step index or count");
  if (cell_begin < 0 || cell_end > p->n_cells || cell_begin >= cell_end || cell_begin % (int32_t)kChunk ||
      cell_end % (int32_t)kChunk)
    return fail(LBM3D_EINVAL, "%s", "cell range: 0 <= cell_begin < cell_end <= n_cells, both multiples of "
                                    "LBM3D_BATCH_ALIGN");
  if (cfg && (!valid_launch(cfg[0]) || !valid_launch(cfg[1])))
    return fail(LBM3D_EINVAL, "%s", "launch shape: block must be 128, 256 or 512 and grid_per_cu 0..64");
  for (int f = 0; f < 6; ++f)
    if (p->bc_mode[f] == 1 && !face_u) return fail(LBM3D_EINVAL, "%s", "face_u is NULL with a pressure face");
  const KP kp = make_kp_batch(*p);
  const Batch bt = batch_view(desc, n_samples, active);
  hipStream_t s = (hipStream_t)stream;
  const uint4 *r = (const uint4 *)rows;
  const uint32_t cb = (uint32_t)cell_begin, ce = (uint32_t)cell_end;
  for (int32_t k = 0; k < nsteps; ++k) {
    const int64_t t = t0 + k;
    switch (dtype) {
      case LBM3D_F32: rc = launch_batch_step<float>(s, kp, bt, g, cells, r, face_bits, face_u, t, cfg, cb, ce); break;
      case LBM3D_F16: rc = launch_batch_step<__half>(s, kp, bt, g, cells, r, face_bits, face_u, t, cfg, cb, ce); break;
      case LBM3D_F64: rc = launch_batch_step<double>(s, kp, bt, g, cells, r, face_bits, face_u, t, cfg, cb, ce); break;
      default:
        rc = launch_batch_step<__hip_bfloat16>(s, kp, bt, g, cells, r, face_bits, face_u, t, cfg, cb, ce);
        break;
    }
    if (rc) return rc;
  }
  return LBM3D_OK;
}

int lbm3d_batch_step(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc, int32_t n_samples,
                     const int32_t *cells, const lbm3d_row_word *rows, const uint32_t *face_bits, void *face_u,
                     const uint8_t *active, int64_t t, const lbm3d_launch *cfg) {
  return lbm3d_batch_run(stream, p, dtype, g, desc, n_samples, cells, rows, face_bits, face_u, active, t, 1, cfg);
}

int lbm3d_batch_macros(void *stream, const lbm3d_params *p, int dtype, const void *g, const void *desc,
                       int32_t n_samples, const int32_t *cells, const lbm3d_row_word *rows,
                       const uint32_t *face_bits, const void *face_u, int64_t t, void *rho, void *v) {
  int rc = check_batch_solver(p, dtype, n_samples, desc, g, cells, rows, face_bits);
  if (rc) return rc;
  if (t < 0) return fail(LBM3D_EINVAL, "%s", "negative step index");
  if (!rho || !v) return fail(LBM3D_EINVAL, "%s", "rho or v output is NULL");
  const KP kp = make_kp_batch(*p);
  const Batch bt = batch_view(desc, n_samples, nullptr);
  hipStream_t s = (hipStream_t)stream;
  const uint4 *r = (const uint4 *)rows;
  switch (dtype) {
    case LBM3D_F32: return launch_batch_macros<float>(s, kp, bt, g, cells, r, face_bits, face_u, t, rho, v);
    case LBM3D_F16: return launch_batch_macros<__half>(s, kp, bt, g, cells, r, face_bits, face_u, t, rho, v);
    case LBM3D_F64: return launch_batch_macros<double>(s, kp, bt, g, cells, r, face_bits, face_u, t, rho, v);
    default: return launch_batch_macros<__hip_bfloat16>(s, kp, bt, g, cells, r, face_bits, face_u, t, rho, v);
  }
}

size_t lbm3d_batch_reduce_workspace_bytes(int64_t cells) {
  const int64_t chunks = cells > 0 ? (cells + kChunk - 1) / kChunk : 1;
  return (size_t)chunks * kRed * sizeof(double);
}

int lbm3d_batch_reduce(void *stream, const lbm3d_params *p, const void *desc, int32_t n_samples, const float *rho,
                       const float *v, double *out, void *workspace, size_t workspace_bytes) {
  int rc = check_batch(p, n_samples, desc);
  if (rc) return rc;
  if (!rho || !v || !out || !workspace) return fail(LBM3D_EINVAL, "%s", "NULL argument to lbm3d_batch_reduce");
  if (workspace_bytes < lbm3d_batch_reduce_workspace_bytes(p->n_cells))
    return fail(LBM3D_EINVAL, "%s", "reduction workspace too small");
  const Batch bt = batch_view(desc, n_samples, nullptr);
  hipStream_t s = (hipStream_t)stream;
  const uint32_t chunks = (uint32_t)p->n_cells / kChunk;
  k_reduce_batch_partial<<<chunks, kBlock, 0, s>>>(bt, rho, v, (uint32_t)p->n_cells, (double *)workspace);
  if ((rc = check_launch("lbm3d_batch_reduce"))) return rc;
  k_reduce_batch_final<<<(unsigned)n_samples, 64 * kRed, 0, s>>>(bt, (const double *)workspace, out);
  return check_launch("lbm3d_batch_reduce/final");
}

}  // extern "C"
