// 3-D "same" convolutions for the warm-start surrogate: BF16 implicit GEMM on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation.  The C
// ABI is declared in include/lbm3d_conv.h.
//
// Forward (and data gradient, which is a forward convolution of the output
// gradient with flipped weights): GEMM with M = voxels, N = output channels,
// K = taps x input channels.  Activations are channels-last, so the 8
// consecutive K values a lane feeds to one MFMA are 8 consecutive channels of
// one (shifted) voxel: a single 16-byte load.  A workgroup of 4 waves computes
// 64 voxels x 16*NT channels; each wave owns 16 voxels.
//
// Weight gradient: GEMM with M = output channels, N = input channels, K =
// voxels.  Channels-last data must be transposed for that K, so each wave
// stages 32 voxels of dy and of the tap-shifted x in LDS as [channel][voxel]
// rows and reads its MFMA operands from there.  Workgroups own (voxel block,
// tap, 16 output channels) and write partial sums; a second kernel folds them
// in a fixed order.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "lbm3d.h"
#include "lbm3d_conv.h"
#include "xcd.h"

#ifndef LBM3D_CONV_TILED
#define LBM3D_CONV_TILED 1  // 3x3x3 forward convolutions stage the input brick in LDS
#endif
#ifndef LBM3D_CONV_XCD_CHUNK
#define LBM3D_CONV_XCD_CHUNK 64  // tiled forward: runs of consecutive bricks per XCD (shared halos hit L2)
#endif
#ifndef LBM3D_CONV_WLDS
#define LBM3D_CONV_WLDS 1  // tiled forward: the chunk's weights staged in LDS (0: loaded per K step)
#endif
#ifndef LBM3D_CONV_WGRAD_TILED
#define LBM3D_CONV_WGRAD_TILED 1  // weight gradients stage dy and x (with its halo) transposed in LDS
#endif
#ifndef LBM3D_CONV_WGRAD_PREFETCH
#define LBM3D_CONV_WGRAD_PREFETCH 0  // 1: load the next brick while the current one's MFMAs run (3 waves/SIMD)
#endif
#ifndef LBM3D_CONV_MW
#define LBM3D_CONV_MW 2  // 16-voxel fragments per wave in the forward kernel
#endif

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

thread_local char c_err[256] = "";

int cfail(int code, const char *msg) {
  snprintf(c_err, sizeof(c_err), "%s", msg);
  return code;
}

int ccheck(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(c_err, sizeof(c_err), "%s: %s", what, hipGetErrorString(e));
    return LBM3D_EHIP;
  }
  return LBM3D_OK;
}

// exact c / d for c < 2^31: (c * m) >> s
struct Div {
  uint32_t d, m, s;
};
Div make_div(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  Div r;
  r.d = d;
  r.m = (uint32_t)(((1ull << (31 + l)) + d - 1) / d);
  r.s = 31 + l;
  return r;
}
__device__ __forceinline__ uint32_t qdiv(uint32_t c, const Div &v) { return (uint32_t)(((uint64_t)c * v.m) >> v.s); }

struct Vol {
  int32_t n, d, h, w;
  Div dw, dh, dd;
  uint32_t vox;      // n*d*h*w
  int32_t periodic;  // 1: taps past a face wrap to the opposite face; 0: zero padding
};

Vol make_vol(int32_t n, int32_t d, int32_t h, int32_t w, int32_t periodic = 0) {
  Vol v;
  v.periodic = periodic;
  v.n = n;
  v.d = d;
  v.h = h;
  v.w = w;
  v.dw = make_div((uint32_t)w);
  v.dh = make_div((uint32_t)h);
  v.dd = make_div((uint32_t)d);
  v.vox = (uint32_t)((int64_t)n * d * h * w);
  return v;
}

// voxel index -> (b, z, y, x)
__device__ __forceinline__ void decode(const Vol &v, uint32_t m, int &b, int &z, int &y, int &x) {
  const uint32_t t = qdiv(m, v.dw);
  x = (int)(m - t * (uint32_t)v.w);
  const uint32_t t2 = qdiv(t, v.dh);
  y = (int)(t - t2 * (uint32_t)v.h);
  const uint32_t bb = qdiv(t2, v.dd);
  z = (int)(t2 - bb * (uint32_t)v.d);
  b = (int)bb;
}

// tap -> offset (dz, dy, dx); ks = 1 has the single tap 0 at offset 0
__device__ __forceinline__ void tap_offset(int tap, int ks, int &dz, int &dy, int &dx) {
  if (ks == 3) {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
  } else {
    dz = dy = dx = 0;
  }
}

// a coordinate one step outside [0, n) wrapped back in
__device__ __forceinline__ int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// neighbour voxel of (b,z,y,x) at tap offset (|offset| <= 1), wrapped when the
// volume is periodic, else -1 outside the volume (zero padding)
__device__ __forceinline__ int64_t shifted(const Vol &v, int b, int z, int y, int x, int dz, int dy, int dx) {
  int z2 = z + dz, y2 = y + dy, x2 = x + dx;
  if (v.periodic) {
    z2 = wrap1(z2, v.d);
    y2 = wrap1(y2, v.h);
    x2 = wrap1(x2, v.w);
  } else if (z2 < 0 || z2 >= v.d || y2 < 0 || y2 >= v.h || x2 < 0 || x2 >= v.w) {
    return -1;
  }
  return (((int64_t)b * v.d + z2) * v.h + y2) * v.w + x2;
}

__device__ __forceinline__ u32x4 ld16(const uint16_t *p) { return *reinterpret_cast<const u32x4 *>(p); }

__device__ __forceinline__ uint16_t to_bf16(float f) {
  const __hip_bfloat16 h = __float2bfloat16(f);  // round to nearest even
  return *reinterpret_cast<const uint16_t *>(&h);
}

// ---------------------------------------------------------------------------
// Forward: y[v][co] = act(sum_{tap, ci} x[v + off(tap)][ci] * w[co][tap*cin + ci] + bias[co])
// MFMA 16x16x32 lane layout: A (16 x 32): lane l holds row l%16, k = 8*(l/16) .. +7;
// B (32 x 16): column l%16, the same k; C (16 x 16): column l%16, rows 4*(l/16) .. +3.
// ---------------------------------------------------------------------------
// Each wave computes MW fragments of 16 voxels x NT fragments of 16 output
// channels; the next K chunk's operands are loaded while the current chunk's
// MFMAs run.
template <int NT, int MW>
__global__ void __launch_bounds__(256)
    k_conv_fwd(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w, const float *__restrict__ bias,
               uint16_t *__restrict__ y, const Vol vol, int cin, int cout, int ks, int kpad, int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const uint32_t mb = blockIdx.x * (64u * MW) + (uint32_t)wave * (16u * MW);
  const int co0 = blockIdx.y * 16 * NT;
  bool mvalid[MW];
  int vb[MW], vz[MW], vy[MW], vx[MW];
#pragma unroll
  for (int f = 0; f < MW; ++f) {
    const uint32_t m = mb + 16u * f + (uint32_t)r16;
    mvalid[f] = m < vol.vox;
    vb[f] = vz[f] = vy[f] = vx[f] = 0;
    if (mvalid[f]) decode(vol, m, vb[f], vz[f], vy[f], vx[f]);
  }
  const int K = ks * ks * ks * cin;
  const uint16_t *wrow[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) wrow[t] = w + (size_t)(co0 + 16 * t + r16) * kpad + 8 * g;
  f32x4 acc[MW][NT];
#pragma unroll
  for (int f = 0; f < MW; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  // this lane's 8 K values of a chunk: k = kc + 8g .. +7, i.e. channels ci .. ci+7 of tap
  int kk = 8 * g;
  int tap = kk / cin, ci = kk - tap * cin;
  auto load_a = [&](int f, int tp, int c, int k) -> u32x4 {
    u32x4 a = u32x4{0u, 0u, 0u, 0u};
    if (mvalid[f] && k < K) {
      int dz, dy, dx;
      tap_offset(tp, ks, dz, dy, dx);
      const int64_t src = shifted(vol, vb[f], vz[f], vy[f], vx[f], dz, dy, dx);
      if (src >= 0) a = ld16(x + src * cin + c);
    }
    return a;
  };
  u32x4 a[MW], b[NT];
#pragma unroll
  for (int f = 0; f < MW; ++f) a[f] = load_a(f, tap, ci, kk);
#pragma unroll
  for (int t = 0; t < NT; ++t) b[t] = ld16(wrow[t]);
  for (int kc = 0; kc < kpad; kc += 32) {
    int kk2 = kk + 32, ci2 = ci + 32, tap2 = tap;
    while (ci2 >= cin) {
      ci2 -= cin;
      ++tap2;
    }
    u32x4 an[MW], bn[NT];
    const bool more = kc + 32 < kpad;
#pragma unroll
    for (int f = 0; f < MW; ++f) an[f] = more ? load_a(f, tap2, ci2, kk2) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int t = 0; t < NT; ++t) bn[t] = more ? ld16(wrow[t] + kc + 32) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8 bv = __builtin_bit_cast(bf16x8, b[t]);
#pragma unroll
      for (int f = 0; f < MW; ++f)
        acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[f]), bv, acc[f][t], 0, 0, 0);
    }
#pragma unroll
    for (int f = 0; f < MW; ++f) a[f] = an[f];
#pragma unroll
    for (int t = 0; t < NT; ++t) b[t] = bn[t];
    kk = kk2;
    ci = ci2;
    tap = tap2;
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = co0 + 16 * t + r16;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int f = 0; f < MW; ++f)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t mo = mb + 16u * f + (uint32_t)(4 * g + e);
        if (mo >= vol.vox) continue;
        float val = acc[f][t][e] + bb;
        if (relu) val = val > 0.f ? val : 0.f;
        y[(size_t)mo * cout + co] = to_bf16(val);
      }
  }
}

// ---------------------------------------------------------------------------
// Forward 3x3x3 convolution with the input tile in LDS.  A workgroup owns a
// brick of 2 x 8 x 16 output voxels (z, y, x) of one volume and 16 output
// channels (blockIdx.y).  For each chunk of 16 input channels it stages the
// brick's halo (4 x 10 x 18 voxels) and the chunk's weight fragments of all 14
// K steps in LDS in one round trip, so the K loop itself only reads LDS and
// issues MFMAs.  One K step covers two taps x
// 16 channels: lane group g takes tap 2s + g/2, channels 8*(g%2) .. +7.  Each
// wave computes 64 of the brick's voxels (4 fragments of 16).
// ---------------------------------------------------------------------------
constexpr int kBz = 2, kBy = 8, kBx = 16;                  // brick
constexpr int kHz = kBz + 2, kHy = kBy + 2, kHx = kBx + 2;  // halo
constexpr int kHalo = kHz * kHy * kHx;                     // 720 voxels
constexpr int kSteps = 14;                                 // K steps per 16-channel chunk (28 tap slots)

// halo-index offset of tap t (dz, dy, dx in -1..1)
__host__ __device__ constexpr int tap_halo_offset(int t) {
  return (t / 9 - 1) * kHy * kHx + ((t / 3) % 3 - 1) * kHx + (t % 3 - 1);
}

__global__ void __launch_bounds__(256)
    k_conv_fwd_tiled(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                     const float *__restrict__ bias, uint16_t *__restrict__ y, const Vol vol, int cin, int cout,
                     int kpad, int relu, int nbz, int nby, int nbx, uint32_t nbricks) {
  __shared__ u32x4 halo[kHalo * 2];  // [voxel][16 channels] as two 16-byte halves
#if LBM3D_CONV_WLDS
  __shared__ u32x4 wl[kSteps][64];    // [K step][lane]: the lane's B fragment
#endif
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int half = g & 1, second = g >> 1;  // channel half and tap parity of this lane's K values
  // brick coordinates (bricks x fastest, then y, z, volume)
  uint32_t bi = LBM3D_CONV_XCD_CHUNK > 0
                    ? xcd_block(LBM3D_CONV_XCD_CHUNK, blockIdx.x & 7u, blockIdx.x >> 3, nbricks)
                    : blockIdx.x;
  if (bi >= nbricks) return;  // padding block of the grid
  const int bx = (int)(bi % (uint32_t)nbx);
  bi /= (uint32_t)nbx;
  const int by = (int)(bi % (uint32_t)nby);
  bi /= (uint32_t)nby;
  const int bz = (int)(bi % (uint32_t)nbz);
  const int bb = (int)(bi / (uint32_t)nbz);
  const int z0 = bz * kBz, y0 = by * kBy, x0 = bx * kBx;
  const int co = blockIdx.y * 16 + r16;  // this lane's output channel in the B and C fragments
  const int64_t vstride = (int64_t)vol.d * vol.h * vol.w;  // voxels per volume
  const uint16_t *xv = x + (size_t)bb * vstride * cin;     // this brick's volume
  // The A operand of voxel fragment f, tap t: halo entry 2 * (hc_f + T(t)) + half,
  // hc_f = the halo index of this lane's voxel (brick-local 64*wave + 16*f + r16:
  // x = r16, y = 4*(wave&1) + f, z = wave>>1) and T(t) the tap's halo offset.  As
  // a byte address: ab_f + 32 T(t) with ab_f = 16 (2 hc_f + half).
  uint32_t ab[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int hc = (((wave >> 1) + 1) * kHy + (4 * (wave & 1) + f + 1)) * kHx + (r16 + 1);
    ab[f] = (uint32_t)(16 * (2 * hc + half));
  }
  f32x4 acc[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
  const char *hb = reinterpret_cast<const char *>(halo);
  for (int c0 = 0; c0 < cin; c0 += 16) {
    __syncthreads();  // the previous chunk's LDS reads are done
#if LBM3D_CONV_WLDS
    for (int e = threadIdx.x; e < kSteps * 64; e += 256) {
      const int st = e >> 6, ln = e & 63, lg = ln >> 4;
      const int tap = 2 * st + (lg >> 1);
      wl[st][ln] = tap < 27 ? ld16(w + (size_t)(blockIdx.y * 16 + (ln & 15)) * kpad + tap * cin + c0 + 8 * (lg & 1))
                            : u32x4{0u, 0u, 0u, 0u};
    }
#endif
    // halo: one voxel (its 16 channels, two 16-byte loads) per thread and pass
    for (int hv = threadIdx.x; hv < kHalo; hv += 256) {
      const int hz = hv / (kHy * kHx), r = hv - hz * (kHy * kHx);
      const int hy = r / kHx, hx = r - hy * kHx;
      int zz = z0 + hz - 1, yy = y0 + hy - 1, xx = x0 + hx - 1;
      if (vol.periodic) {  // halo voxels past a face come from the opposite face
        zz = wrap1(zz, vol.d);
        yy = wrap1(yy, vol.h);
        xx = wrap1(xx, vol.w);
      }
      u32x4 lo = u32x4{0u, 0u, 0u, 0u}, hi = lo;
      if ((unsigned)zz < (unsigned)vol.d && (unsigned)yy < (unsigned)vol.h && (unsigned)xx < (unsigned)vol.w) {
        const uint16_t *src = xv + (((int64_t)zz * vol.h + yy) * vol.w + xx) * cin + c0;
        lo = ld16(src);
        hi = ld16(src + 8);
      }
      halo[2 * hv] = lo;
      halo[2 * hv + 1] = hi;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
      // taps 2s and 2s+1 (this lane: 2s + second); the 28th slot (s = 13,
      // second half) is zero padding
      constexpr int kMin = -(kHy * kHx + kHx + 1);  // smallest tap offset
      const int t1 = 2 * s + 1 < 27 ? 2 * s + 1 : 26;
      const uint32_t dsec = second ? (uint32_t)(32 * (tap_halo_offset(t1) - tap_halo_offset(2 * s))) : 0u;
      const uint32_t imm = (uint32_t)(32 * (tap_halo_offset(2 * s) - kMin));
#if LBM3D_CONV_WLDS
      const bf16x8 bv = __builtin_bit_cast(bf16x8, wl[s][lane]);
#else
      const bf16x8 bv = __builtin_bit_cast(
          bf16x8, 2 * s + second < 27 ? ld16(w + (size_t)co * kpad + (2 * s + second) * cin + c0 + 8 * half)
                                      : u32x4{0u, 0u, 0u, 0u});
#endif
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        u32x4 a = *reinterpret_cast<const u32x4 *>(hb + (ab[f] + (uint32_t)(32 * kMin) + dsec) + imm);
        if (s == kSteps - 1 && second) a = u32x4{0u, 0u, 0u, 0u};
        acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), bv, acc[f], 0, 0, 0);
      }
    }
  }
  // C layout: lane holds rows (voxels) 4g .. 4g+3 of each fragment, column
  // (channel) r16: voxel x = 4g + e, y = 4*(wave&1) + f, z = wave>>1 of the brick
  const float bsum = bias ? bias[co] : 0.f;
  const int oz = z0 + (wave >> 1), oy0 = y0 + 4 * (wave & 1), ox0 = x0 + 4 * g;
  if (oz >= vol.d) return;
  uint16_t *yv = y + (size_t)bb * vstride * cout + (((int64_t)oz * vol.h + oy0) * vol.w + ox0) * cout + co;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (oy0 + f >= vol.h) break;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (ox0 + e >= vol.w) break;
      float val = acc[f][e] + bsum;
      if (relu) val = val > 0.f ? val : 0.f;
      yv[((int64_t)f * vol.w + e) * cout] = to_bf16(val);
    }
  }
}

// wt[ci][tap'*cout + co] = w[co][tap*cin + ci], tap' = taps-1-tap; zero padding columns
__global__ void k_wflip(const uint16_t *__restrict__ w, uint16_t *__restrict__ wt, int cin, int cout, int taps,
                        int kpad_in, int kpad_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)cin * kpad_out;
  if (i >= total) return;
  const int ci = (int)(i / kpad_out), k = (int)(i % kpad_out);
  uint16_t val = 0;
  if (k < taps * cout) {
    const int tp = k / cout, co = k % cout;
    const int tap = taps - 1 - tp;
    val = w[(size_t)co * kpad_in + tap * cin + ci];
  }
  wt[i] = val;
}

// ---------------------------------------------------------------------------
// Weight gradient.  Workgroup (p, tap, cot): voxels [p*vblock, (p+1)*vblock),
// one tap, output channels 16*cot .. +15; 2 waves, each taking 32 of every 64
// voxels.  part[((p*2 + wave)*taps + tap)*cout + co][ci].
// ---------------------------------------------------------------------------
constexpr int kWgWaves = 2;
constexpr int kLd = 40;        // LDS row stride (bf16): 80 bytes, conflict-free 16-byte reads
constexpr int kMaxCinTiles = 16;  // cin <= 256

template <int MAXT>
__global__ void __launch_bounds__(64 * kWgWaves)
    k_conv_wgrad(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dy, float *__restrict__ part,
                 const Vol vol, int cin, int cout, int ks, uint32_t vblock) {
  extern __shared__ uint16_t lds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int p = blockIdx.x, tap = blockIdx.y, co0 = blockIdx.z * 16;
  const int taps = ks * ks * ks;
  const int nt = cin / 16;
  uint16_t *dyT = lds + (size_t)wave * (16 + cin) * kLd;  // [16][kLd]
  uint16_t *xT = dyT + 16 * kLd;                           // [cin][kLd]
  int dz, dyo, dx;
  tap_offset(tap, ks, dz, dyo, dx);
  const uint32_t v_begin = (uint32_t)p * vblock;
  uint32_t v_end = v_begin + vblock;
  if (v_end > vol.vox) v_end = vol.vox;
  f32x4 acc[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int groups = cin / 8;            // 8-channel groups per voxel row
  const int pieces = 32 * groups;        // 16-byte pieces of x per 32 voxels
  for (uint32_t base = v_begin; base < v_end; base += 32u * kWgWaves) {
    const uint32_t c0 = base + 32u * (uint32_t)wave;  // this wave's 32 voxels
    // dy^T: lane -> voxel c0 + lane/2, channels co0 + 8*(lane&1) .. +7
    {
      const uint32_t v = c0 + (uint32_t)(lane >> 1);
      u32x4 val = u32x4{0u, 0u, 0u, 0u};
      if (v < v_end) val = ld16(dy + (size_t)v * cout + co0 + 8 * (lane & 1));
      const uint16_t *h = reinterpret_cast<const uint16_t *>(&val);
#pragma unroll
      for (int e = 0; e < 8; ++e) dyT[(8 * (lane & 1) + e) * kLd + (lane >> 1)] = h[e];
    }
    // x^T of the tap-shifted voxels: piece -> voxel piece % 32, channel group piece / 32
    for (int pc = lane; pc < pieces; pc += 64) {
      const int vv = pc & 31, grp = pc >> 5;
      const uint32_t v = c0 + (uint32_t)vv;
      u32x4 val = u32x4{0u, 0u, 0u, 0u};
      if (v < v_end) {
        int b, z, yy, xx;
        decode(vol, v, b, z, yy, xx);
        const int64_t src = shifted(vol, b, z, yy, xx, dz, dyo, dx);
        if (src >= 0) val = ld16(x + src * cin + 8 * grp);
      }
      const uint16_t *h = reinterpret_cast<const uint16_t *>(&val);
#pragma unroll
      for (int e = 0; e < 8; ++e) xT[(8 * grp + e) * kLd + vv] = h[e];
    }
    __syncthreads();
    const bf16x8 av = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(dyT + r16 * kLd + 8 * g));
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      if (t < nt) {
        const bf16x8 bv =
            __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(xT + (16 * t + r16) * kLd + 8 * g));
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float *out = part + ((((size_t)p * kWgWaves + wave) * taps + tap) * cout + co0) * cin;
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    if (t < nt) {
#pragma unroll
      for (int e = 0; e < 4; ++e) out[(size_t)(4 * g + e) * cin + 16 * t + r16] = acc[t][e];
    }
  }
}

// dw[co][tap*cin + ci] = sum over the partial sets, in order; padding columns 0
// ---------------------------------------------------------------------------
// Weight gradient with the operands in LDS:
//   dw[co][tap*cin + ci] = sum over voxels v of dy[v][co] * x[v + off(tap)][ci].
// A workgroup owns 16 output channels co0.., 16 input channels c0.. and a run
// of bricks (2 x 8 x 16 voxels, as the forward kernel's).  Per brick it stages
// dy transposed, dyT[co][voxel], and x transposed, xs[ci][...]: for 3x3x3
// kernels x's halo as [ci][halo row][x], for 1x1x1 kernels the brick itself as
// [ci][voxel].  For each K chunk of 32 brick voxels (two brick rows) a wave
// reads one A fragment from dyT and, for each of its taps, one B fragment from
// xs: the 8 halo voxels the chunk's voxels see through that tap, consecutive
// in an xs row but offset by dx + 1 elements, so assembled from one or two
// aligned LDS reads.  3x3x3: wave w accumulates taps w, w+4, ... over all
// chunks.  1x1x1: wave w accumulates chunks 2w and 2w+1.  Each wave writes one
// partial 16 x 16 block per tap; k_wgrad_fold sums the partials in a fixed
// order.
// ---------------------------------------------------------------------------
constexpr int kXRow = 24;  // halo row in xs, bf16 (18 used; 48 bytes keep rows 16-byte aligned)
constexpr int kXRows = kHz * kHy;  // 40 halo rows
// channel strides of the LDS tiles, in bf16: 4 dwords modulo 32 banks, so the
// 16 lanes reading 16 channels at one position hit distinct banks
constexpr int kXChan = kXRows * kXRow + 8;  // 968
constexpr int kDyRow = 256 + 8;             // 264
constexpr int wgrad_sets(int ks) { return ks == 3 ? 1 : 4; }  // partial sets a workgroup writes

template <int KS>
__global__ void __launch_bounds__(256)
    k_conv_wgrad_tiled(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dy, float *__restrict__ part,
                       const Vol vol, int cin, int cout, int nbz, int nby, int nbx, uint32_t nbricks,
                       uint32_t per_block) {
  constexpr int taps = KS * KS * KS;
  constexpr int xchan = KS == 3 ? kXChan : kDyRow;
  constexpr int nacc = KS == 3 ? 7 : 1;
  __shared__ __align__(16) uint16_t dyT[16 * kDyRow];
  __shared__ __align__(16) uint16_t xs[16 * xchan];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int nci = cin / 16;
  const int co0 = (blockIdx.y / nci) * 16, c0 = (blockIdx.y % nci) * 16;
  const uint32_t b0 = blockIdx.x * per_block;
  const uint32_t b1 = b0 + per_block < nbricks ? b0 + per_block : nbricks;
  const int64_t vstride = (int64_t)vol.d * vol.h * vol.w;
  f32x4 acc[nacc];
#pragma unroll
  for (int i = 0; i < nacc; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Global loads of one brick into registers: dy (and x for 1x1x1) of this
  // thread's brick voxel, and for 3x3x3 up to three of x's halo voxels.
  constexpr int nh = KS == 3 ? (kHalo + 255) / 256 : 0;
  u32x4 rdy[2], rx[2], rh[nh > 0 ? nh : 1][2];
  auto load_brick = [&](uint32_t br) {
    uint32_t bi = br;
    const int bx = (int)(bi % (uint32_t)nbx);
    bi /= (uint32_t)nbx;
    const int by = (int)(bi % (uint32_t)nby);
    bi /= (uint32_t)nby;
    const int bz = (int)(bi % (uint32_t)nbz);
    const int bb = (int)(bi / (uint32_t)nbz);
    const int z0 = bz * kBz, y0 = by * kBy, x0 = bx * kBx;
    const u32x4 zero = u32x4{0u, 0u, 0u, 0u};
    {
      const int t = threadIdx.x;
      const int z = z0 + (t >> 7), y = y0 + ((t >> 4) & 7), xx = x0 + (t & 15);
      rdy[0] = rdy[1] = rx[0] = rx[1] = zero;
      if (z < vol.d && y < vol.h && xx < vol.w) {
        const size_t v = (size_t)bb * vstride + ((int64_t)z * vol.h + y) * vol.w + xx;
        rdy[0] = ld16(dy + v * cout + co0);
        rdy[1] = ld16(dy + v * cout + co0 + 8);
        if (KS == 1) {
          rx[0] = ld16(x + v * cin + c0);
          rx[1] = ld16(x + v * cin + c0 + 8);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < nh; ++q) {
      const int hv = threadIdx.x + 256 * q;
      rh[q][0] = rh[q][1] = zero;
      const int hz = hv / (kHy * kHx), r = hv - hz * (kHy * kHx);
      const int hy = r / kHx, hx = r - hy * kHx;
      const int zz = z0 + hz - 1, yy = y0 + hy - 1, xx = x0 + hx - 1;
      if (hv < kHalo && (unsigned)zz < (unsigned)vol.d && (unsigned)yy < (unsigned)vol.h &&
          (unsigned)xx < (unsigned)vol.w) {
        const uint16_t *src = x + ((size_t)bb * vstride + ((int64_t)zz * vol.h + yy) * vol.w + xx) * cin + c0;
        rh[q][0] = ld16(src);
        rh[q][1] = ld16(src + 8);
      }
    }
  };
  // registers -> LDS, transposed
  auto store_brick = [&]() {
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      dyT[c * kDyRow + t] = (uint16_t)(rdy[0][c >> 1] >> (16 * (c & 1)));
      dyT[(8 + c) * kDyRow + t] = (uint16_t)(rdy[1][c >> 1] >> (16 * (c & 1)));
      if (KS == 1) {
        xs[c * xchan + t] = (uint16_t)(rx[0][c >> 1] >> (16 * (c & 1)));
        xs[(8 + c) * xchan + t] = (uint16_t)(rx[1][c >> 1] >> (16 * (c & 1)));
      }
    }
#pragma unroll
    for (int q = 0; q < nh; ++q) {
      const int hv = threadIdx.x + 256 * q;
      if (hv < kHalo) {
        const int hz = hv / (kHy * kHx), r = hv - hz * (kHy * kHx);
        const int hy = r / kHx, hx = r - hy * kHx;
        const int e = (hz * kHy + hy) * kXRow + hx;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          xs[c * xchan + e] = (uint16_t)(rh[q][0][c >> 1] >> (16 * (c & 1)));
          xs[(8 + c) * xchan + e] = (uint16_t)(rh[q][1][c >> 1] >> (16 * (c & 1)));
        }
      }
    }
  };
  if (LBM3D_CONV_WGRAD_PREFETCH) load_brick(b0);
  for (uint32_t br = b0; br < b1; ++br) {
    if (!LBM3D_CONV_WGRAD_PREFETCH) load_brick(br);
    __syncthreads();  // the previous brick's LDS reads are done
    store_brick();
    __syncthreads();
    // the next brick's loads are in flight while this one's MFMAs run
    if (LBM3D_CONV_WGRAD_PREFETCH && br + 1 < b1) load_brick(br + 1);
    if (KS == 1) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int k0 = 32 * (2 * wave + q) + 8 * g;
        const u32x4 a = *reinterpret_cast<const u32x4 *>(&dyT[r16 * kDyRow + k0]);
        const u32x4 b = *reinterpret_cast<const u32x4 *>(&xs[r16 * xchan + k0]);
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                         acc[0], 0, 0, 0);
      }
    } else {
#pragma unroll 2
      for (int kc = 0; kc < 8; ++kc) {
        // K values 8g .. 8g+7 of this chunk: brick voxels x = 8(g&1) .. +7 of
        // row y = (2kc + g/2) % 8, z = kc / 4
        const u32x4 a = *reinterpret_cast<const u32x4 *>(&dyT[r16 * kDyRow + 32 * kc + 8 * g]);
        const int lz = kc >> 2, ly = (2 * kc + (g >> 1)) & 7, jx = 8 * (g & 1);
#pragma unroll
        for (int i = 0; i < nacc; ++i) {
          const int t = wave + 4 * i;  // wave-uniform
          if (t < taps) {
            const int dz = t / 9 - 1, dyo = (t / 3) % 3 - 1, dx = t % 3 - 1;
            const uint16_t *rowp = &xs[r16 * xchan + ((lz + dz + 1) * kHy + (ly + dyo + 1)) * kXRow + jx];
            // halo x = jx + dx + 1 .. +7, shifted by dx + 1 elements from an aligned read
            const u32x4 w0 = *reinterpret_cast<const u32x4 *>(rowp);
            u32x4 b = w0;
            if (dx >= 0) {
              const uint32_t w4 = *reinterpret_cast<const uint32_t *>(rowp + 8);
              if (dx == 0)
                b = u32x4{__builtin_amdgcn_alignbyte(w0[1], w0[0], 2u), __builtin_amdgcn_alignbyte(w0[2], w0[1], 2u),
                          __builtin_amdgcn_alignbyte(w0[3], w0[2], 2u), __builtin_amdgcn_alignbyte(w4, w0[3], 2u)};
              else
                b = u32x4{w0[1], w0[2], w0[3], w4};
            }
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                             __builtin_bit_cast(bf16x8, b), acc[i], 0, 0, 0);
          }
        }
      }
    }
  }
  // C layout: lane holds rows (co) 4g .. 4g+3, column (ci) r16; part[set][tap][co][ci]
  // with set = block (3x3x3) or 4 * block + wave (1x1x1)
#pragma unroll
  for (int i = 0; i < nacc; ++i) {
    const int t = KS == 3 ? wave + 4 * i : 0;
    const size_t set = KS == 3 ? blockIdx.x : (size_t)blockIdx.x * 4 + wave;
    if (t >= taps) break;
    float *dst = part + (set * taps + t) * cout * cin + (size_t)(co0 + 4 * g) * cin + c0 + r16;
#pragma unroll
    for (int e = 0; e < 4; ++e) dst[(size_t)e * cin] = acc[i][e];
  }
}

// One pass of the tree fold of partial weight-gradient sets, in place: the
// thread for (group q, element e) sums sets q*16*stride + j*stride, j < 16, at
// element e into the group's first set.  Passes with stride 1, 16, 256, ...
// leave the total in set 0, summed in an order fixed by the set count alone.
__global__ void k_wgrad_fold_pass(float *__restrict__ part, int64_t set_size, int nsets, int stride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t groups = (nsets + 16ll * stride - 1) / (16ll * stride);
  if (i >= groups * set_size) return;
  const int64_t q = i / set_size, e = i - q * set_size;
  const int64_t first = q * 16 * stride;
  float v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int64_t set = first + (int64_t)j * stride;
    v[j] = set < nsets ? part[set * set_size + e] : 0.f;
  }
#pragma unroll
  for (int w = 8; w >= 1; w >>= 1)
#pragma unroll
    for (int j = 0; j < w; ++j) v[j] += v[j + w];
  part[first * set_size + e] = v[0];
}

// Sums nsets partial sets into dw: tree passes of k_wgrad_fold_pass while more
// than 16 sets remain, then k_wgrad_fold over the survivors.
int wgrad_fold(hipStream_t s, float *part, float *dw, int nsets, int taps, int cin, int cout, int kpad);

__global__ void k_wgrad_fold(const float *__restrict__ part, float *__restrict__ dw, int nsets, int set_step, int taps,
                             int cin, int cout, int kpad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)cout * kpad) return;
  const int co = (int)(i / kpad), k = (int)(i % kpad);
  float s = 0.f;
  if (k < taps * cin) {
    const int tap = k / cin, ci = k % cin;
    const size_t stride = (size_t)taps * cout * cin;
    const size_t off = ((size_t)tap * cout + co) * cin + ci;
    for (int q = 0; q < nsets; ++q) s += part[(size_t)q * set_step * stride + off];
  }
  dw[i] = s;
}

int wgrad_fold(hipStream_t s, float *part, float *dw, int nsets, int taps, int cin, int cout, int kpad) {
  const int64_t set_size = (int64_t)taps * cout * cin;
  int stride = 1;
  while (nsets > 16 * stride) {  // sets 0, stride, 2*stride, ... hold the partial sums
    const int64_t groups = (nsets + 16ll * stride - 1) / (16ll * stride);
    const int64_t threads = groups * set_size;
    k_wgrad_fold_pass<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(part, set_size, nsets, stride);
    if (int rc = ccheck("lbm3d_conv3d_wgrad/fold pass")) return rc;
    stride *= 16;
  }
  const int64_t total = (int64_t)cout * kpad;
  k_wgrad_fold<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(part, dw, (nsets + stride - 1) / stride, stride, taps,
                                                                cin, cout, kpad);
  return ccheck("lbm3d_conv3d_wgrad/fold");
}

int check_conv(int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t ks) {
  if (n < 1 || d < 1 || h < 1 || wd < 1) return cfail(LBM3D_EINVAL, "volume extents must be >= 1");
  if (ks != 1 && ks != 3) return cfail(LBM3D_EINVAL, "kernel size must be 1 or 3");
  if (cin < 16 || cin > 256 || cin % 16 || cout < 16 || cout > 256 || cout % 16)
    return cfail(LBM3D_EINVAL, "channel counts must be multiples of 16 in [16, 256]");
  // voxel indices are 32-bit (exact magic division below 2^31); element offsets are 64-bit
  if ((int64_t)n * d * h * wd >= ((int64_t)1 << 31)) return cfail(LBM3D_ESIZE, "more than 2^31-1 voxels");
  return LBM3D_OK;
}

constexpr int kWgTargetBlocks = 2048;

// Tiled weight gradient: bricks per workgroup and workgroups along the brick
// range, about kWgTargetBlocks workgroups in all.
int wgrad_tiled_plan(int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout, uint32_t &nbricks,
                     uint32_t &per) {
  const int64_t nb = (int64_t)n * ((d + kBz - 1) / kBz) * ((h + kBy - 1) / kBy) * ((wd + kBx - 1) / kBx);
  const int64_t tiles = (int64_t)(cout / 16) * (cin / 16);
  int64_t P = (kWgTargetBlocks + tiles - 1) / tiles;
  if (P > nb) P = nb;
  if (P < 1) P = 1;
  const int64_t pb = (nb + P - 1) / P;
  nbricks = (uint32_t)nb;
  per = (uint32_t)pb;
  return (int)((nb + pb - 1) / pb);
}

void wgrad_plan(int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cout, int32_t ks, int &parts, uint32_t &vblock) {
  const int64_t vox = (int64_t)n * d * h * wd;
  const int taps = ks * ks * ks;
  const int64_t per_part = (int64_t)taps * (cout / 16);
  int64_t P = (kWgTargetBlocks + per_part - 1) / per_part;
  const int64_t step = 32 * kWgWaves;
  const int64_t max_parts = (vox + step - 1) / step;
  if (P > max_parts) P = max_parts;
  if (P < 1) P = 1;
  int64_t vb = (vox + P - 1) / P;
  vb = (vb + step - 1) / step * step;
  parts = (int)((vox + vb - 1) / vb);
  vblock = (uint32_t)vb;
}

// ---------------------------------------------------------------------------
// Periodic distance to solid (the surrogate's second input channel): the
// separable squared-distance transform as three min-plus passes,
//   g(v) = min over |d| <= M of f(v + d e_axis) + d^2  (coordinates wrapped),
// starting from f = 0 on solid cells and 3M^2 + 1 elsewhere, then
// dist = min(sqrt(f), M).  Exact for distances up to M; larger ones are capped.
// Threads run along the innermost axis, so every pass reads coalesced rows.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_dist_pass(const int8_t *__restrict__ solid, const float *__restrict__ f, float *__restrict__ g,
                int64_t total, int d0, int d1, int d2, int axis, int margin, int last) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= total) return;
  const int z = (int)(v % d2);
  const int64_t r = v / d2;
  const int y = (int)(r % d1);
  const int x = (int)((r / d1) % d0);
  const int n = axis == 0 ? d0 : (axis == 1 ? d1 : d2);
  const int c = axis == 0 ? x : (axis == 1 ? y : z);
  const int64_t stride = axis == 0 ? (int64_t)d1 * d2 : (axis == 1 ? (int64_t)d2 : 1);
  const int64_t line = v - (int64_t)c * stride;  // this voxel's line at coordinate 0
  const float big = (float)(3 * margin * margin + 1);
  float best = 3.0e38f;
  for (int d = -margin; d <= margin; ++d) {
    int cc = (c + d) % n;
    if (cc < 0) cc += n;
    const int64_t u = line + (int64_t)cc * stride;
    const float fv = solid ? (solid[u] != 0 ? 0.f : big) : f[u];
    best = fminf(best, fv + (float)(d * d));
  }
  // the square root in FP64 rounded once to FP32: the distance an exact EDT
  // (FP64) reports, cast to FP32
  g[v] = last ? fminf((float)sqrt((double)best), (float)margin) : best;
}

// bf16 y > 0 (false for -0, negatives and NaN)
__device__ __forceinline__ bool bf16_positive(uint32_t h) {
  return !(h & 0x8000u) && h != 0u && !((h & 0x7F80u) == 0x7F80u && (h & 0x7Fu));
}

__device__ __forceinline__ float bf16_to_f32(uint32_t h) { return __uint_as_float(h << 16); }

// Backward of y = relu(z): dz = g * (y > 0) (a masked element is g * 0.0 in
// FP32, rounded to bf16: a zero with g's sign, or NaN), stored to dz; with
// y == nullptr dz = g.  Each block also sums dz per channel over its voxels
// in FP32: part[block][c].  Data are [voxels][C] bf16, C dividing 2048, so a
// thread's 8-element chunks (chunk q covers elements 8q .. 8q+7) always hold
// the same 8 channels when the blocks' chunk ranges start on multiples of C/8.
__global__ void __launch_bounds__(256)
    k_relu_bias_grad(const uint16_t *g, const uint16_t *__restrict__ y, uint16_t *dz,  // dz may be g
                     float *__restrict__ part, int64_t chunks, int C, int64_t per_block) {
  __shared__ float red[256][9];
  const int t = (int)threadIdx.x;
  const int64_t q0 = (int64_t)blockIdx.x * per_block;
  const int64_t q1 = q0 + per_block < chunks ? q0 + per_block : chunks;
  float acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.f;
  for (int64_t q = q0 + t; q < q1; q += 256) {
    u32x4 a = ld16(g + q * 8);
    if (y) {
      const u32x4 m = ld16(y + q * 8);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        uint32_t lo = a[w] & 0xFFFFu, hi = a[w] >> 16;
        if (!bf16_positive(m[w] & 0xFFFFu))
          lo = __bfloat16_as_ushort(__float2bfloat16(bf16_to_f32(lo) * 0.0f));
        if (!bf16_positive(m[w] >> 16))
          hi = __bfloat16_as_ushort(__float2bfloat16(bf16_to_f32(hi) * 0.0f));
        a[w] = lo | (hi << 16);
      }
    }
    *reinterpret_cast<u32x4 *>(dz + q * 8) = a;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      acc[2 * w] += bf16_to_f32(a[w] & 0xFFFFu);
      acc[2 * w + 1] += bf16_to_f32(a[w] >> 16);
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) red[t][e] = acc[e];
  __syncthreads();
  if (t < C) {  // channel t: threads r, r + C/8, ... hold it in slot t % 8 (r = t / 8)
    const int groups = C / 8;
    float sum = 0.f;
    for (int u = t / 8; u < 256; u += groups) sum += red[u][t % 8];
    part[(int64_t)blockIdx.x * C + t] = sum;
  }
}

int relu_bias_plan(int64_t voxels, int C, int64_t &chunks, int64_t &per_block) {
  chunks = voxels * C / 8;
  int64_t blocks = (chunks + 256 * 8 - 1) / (256 * 8);  // about 8 chunks per thread
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  const int64_t unit = C / 8;
  per_block = (chunks + blocks - 1) / blocks;
  per_block = (per_block + unit - 1) / unit * unit;
  return (int)((chunks + per_block - 1) / per_block);
}

}  // namespace

extern "C" {

const char *lbm3d_conv_last_error(void) { return c_err; }

int32_t lbm3d_conv_kpad(int32_t cin, int32_t ks) {
  if (cin < 1 || (ks != 1 && ks != 3)) return 0;
  const int32_t k = ks * ks * ks * cin;
  return (k + 31) / 32 * 32;
}

int lbm3d_conv3d_fwd(void *stream, const uint16_t *x, const uint16_t *w, const float *bias, uint16_t *y, int32_t n,
                     int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t ks, int32_t flags) {
  int rc = check_conv(n, d, h, wd, cin, cout, ks);
  if (rc) return rc;
  if (!x || !w || !y) return cfail(LBM3D_EINVAL, "NULL tensor");
  if (flags & ~(LBM3D_CONV_RELU | LBM3D_CONV_PERIODIC)) return cfail(LBM3D_EINVAL, "unknown flag bits");
  const int relu = (flags & LBM3D_CONV_RELU) ? 1 : 0;
  const Vol vol = make_vol(n, d, h, wd, (flags & LBM3D_CONV_PERIODIC) ? 1 : 0);
  const int kpad = lbm3d_conv_kpad(cin, ks);
  hipStream_t s = (hipStream_t)stream;
  if (ks == 3 && LBM3D_CONV_TILED) {
    const int nbz = (d + kBz - 1) / kBz, nby = (h + kBy - 1) / kBy, nbx = (wd + kBx - 1) / kBx;
    const int64_t bricks = (int64_t)n * nbz * nby * nbx;
    const int64_t grid = xcd_grid(bricks, LBM3D_CONV_XCD_CHUNK);
    if (grid > (int64_t)(UINT32_MAX / 256u)) return cfail(LBM3D_ESIZE, "too many bricks for one launch");
    k_conv_fwd_tiled<<<dim3((unsigned)grid, cout / 16), 256, 0, s>>>(x, w, bias, y, vol, cin, cout, kpad, relu, nbz,
                                                                   nby, nbx, (uint32_t)bricks);
    return ccheck("lbm3d_conv3d_fwd");
  }
  constexpr int MW = LBM3D_CONV_MW;
  const unsigned gx = (vol.vox + 64u * MW - 1u) / (64u * MW);
  if (cout % 64 == 0)
    k_conv_fwd<4, MW><<<dim3(gx, cout / 64), 256, 0, s>>>(x, w, bias, y, vol, cin, cout, ks, kpad, relu);
  else if (cout % 32 == 0)
    k_conv_fwd<2, MW><<<dim3(gx, cout / 32), 256, 0, s>>>(x, w, bias, y, vol, cin, cout, ks, kpad, relu);
  else
    k_conv_fwd<1, MW><<<dim3(gx, cout / 16), 256, 0, s>>>(x, w, bias, y, vol, cin, cout, ks, kpad, relu);
  return ccheck("lbm3d_conv3d_fwd");
}

int lbm3d_conv3d_wflip(void *stream, const uint16_t *w, uint16_t *wt, int32_t cin, int32_t cout, int32_t ks) {
  int rc = check_conv(1, 1, 1, 1, cin, cout, ks);
  if (rc) return rc;
  if (!w || !wt) return cfail(LBM3D_EINVAL, "NULL tensor");
  const int taps = ks * ks * ks;
  const int kin = lbm3d_conv_kpad(cin, ks), kout = lbm3d_conv_kpad(cout, ks);
  const int64_t total = (int64_t)cin * kout;
  k_wflip<<<(unsigned)((total + 255) / 256), 256, 0, (hipStream_t)stream>>>(w, wt, cin, cout, taps, kin, kout);
  return ccheck("lbm3d_conv3d_wflip");
}

int lbm3d_conv3d_dgrad(void *stream, const uint16_t *dy, const uint16_t *wt, uint16_t *dx, int32_t n, int32_t d,
                       int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t ks) {
  // a forward convolution from cout to cin channels with the flipped weights
  return lbm3d_conv3d_fwd(stream, dy, wt, nullptr, dx, n, d, h, wd, cout, cin, ks, 0);
}

size_t lbm3d_conv3d_wgrad_workspace_bytes(int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout,
                                          int32_t ks) {
  if (check_conv(n, d, h, wd, cin, cout, ks)) return 0;
  if (LBM3D_CONV_WGRAD_TILED) {
    uint32_t nbricks, per;
    const int parts = wgrad_tiled_plan(n, d, h, wd, cin, cout, nbricks, per);
    return (size_t)parts * wgrad_sets(ks) * ks * ks * ks * cout * cin * sizeof(float);
  }
  int parts;
  uint32_t vb;
  wgrad_plan(n, d, h, wd, cout, ks, parts, vb);
  return (size_t)parts * kWgWaves * ks * ks * ks * cout * cin * sizeof(float);
}

int lbm3d_conv3d_wgrad(void *stream, const uint16_t *x, const uint16_t *dy, float *dw, int32_t n, int32_t d, int32_t h,
                       int32_t wd, int32_t cin, int32_t cout, int32_t ks, void *workspace, size_t workspace_bytes) {
  int rc = check_conv(n, d, h, wd, cin, cout, ks);
  if (rc) return rc;
  if (!x || !dy || !dw || !workspace) return cfail(LBM3D_EINVAL, "NULL tensor or workspace");
  if (workspace_bytes < lbm3d_conv3d_wgrad_workspace_bytes(n, d, h, wd, cin, cout, ks))
    return cfail(LBM3D_EINVAL, "wgrad workspace too small");
  const Vol vol = make_vol(n, d, h, wd);
  hipStream_t s = (hipStream_t)stream;
  if (LBM3D_CONV_WGRAD_TILED) {
    uint32_t nbricks, per;
    const int parts = wgrad_tiled_plan(n, d, h, wd, cin, cout, nbricks, per);
    const int nbz = (d + kBz - 1) / kBz, nby = (h + kBy - 1) / kBy, nbx = (wd + kBx - 1) / kBx;
    const dim3 grid((unsigned)parts, (unsigned)((cout / 16) * (cin / 16)));
    if (ks == 3)
      k_conv_wgrad_tiled<3>
          <<<grid, 256, 0, s>>>(x, dy, (float *)workspace, vol, cin, cout, nbz, nby, nbx, nbricks, per);
    else
      k_conv_wgrad_tiled<1>
          <<<grid, 256, 0, s>>>(x, dy, (float *)workspace, vol, cin, cout, nbz, nby, nbx, nbricks, per);
    if ((rc = ccheck("lbm3d_conv3d_wgrad"))) return rc;
    const int kpad = lbm3d_conv_kpad(cin, ks);
    return wgrad_fold(s, (float *)workspace, dw, parts * wgrad_sets(ks), ks * ks * ks, cin, cout, kpad);
  }
  int parts;
  uint32_t vb;
  wgrad_plan(n, d, h, wd, cout, ks, parts, vb);
  const int taps = ks * ks * ks;
  const size_t lds = (size_t)kWgWaves * (16 + cin) * kLd * sizeof(uint16_t);
  const dim3 grid((unsigned)parts, (unsigned)taps, (unsigned)(cout / 16));
  float *part = (float *)workspace;
  if (cin <= 64)
    k_conv_wgrad<4><<<grid, 64 * kWgWaves, lds, s>>>(x, dy, part, vol, cin, cout, ks, vb);
  else if (cin <= 128)
    k_conv_wgrad<8><<<grid, 64 * kWgWaves, lds, s>>>(x, dy, part, vol, cin, cout, ks, vb);
  else
    k_conv_wgrad<kMaxCinTiles><<<grid, 64 * kWgWaves, lds, s>>>(x, dy, part, vol, cin, cout, ks, vb);
  if ((rc = ccheck("lbm3d_conv3d_wgrad"))) return rc;
  const int kpad = lbm3d_conv_kpad(cin, ks);
  const int64_t total = (int64_t)cout * kpad;
  k_wgrad_fold<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(part, dw, parts * kWgWaves, 1, taps, cin, cout, kpad);
  return ccheck("lbm3d_conv3d_wgrad/fold");
}

size_t lbm3d_relu_bias_grad_workspace_bytes(int64_t voxels, int32_t c) {
  if (voxels < 1 || c < 8 || c > 256 || 2048 % c) return 0;
  int64_t chunks, per;
  return (size_t)relu_bias_plan(voxels, c, chunks, per) * c * sizeof(float);
}

int lbm3d_relu_bias_grad(void *stream, const uint16_t *g, const uint16_t *y, uint16_t *dz, float *db, int64_t voxels,
                         int32_t c, void *workspace, size_t workspace_bytes) {
  if (voxels < 1) return cfail(LBM3D_EINVAL, "no voxels");
  if (c < 8 || c > 256 || 2048 % c) return cfail(LBM3D_EINVAL, "channels must divide 2048, 8..256");
  if (!g || !dz || !db || !workspace) return cfail(LBM3D_EINVAL, "NULL tensor or workspace");
  if (workspace_bytes < lbm3d_relu_bias_grad_workspace_bytes(voxels, c))
    return cfail(LBM3D_EINVAL, "relu/bias workspace too small");
  int64_t chunks, per;
  const int blocks = relu_bias_plan(voxels, c, chunks, per);
  hipStream_t s = (hipStream_t)stream;
  k_relu_bias_grad<<<(unsigned)blocks, 256, 0, s>>>(g, y, dz, (float *)workspace, chunks, c, per);
  if (int rc = ccheck("lbm3d_relu_bias_grad")) return rc;
  // the wgrad fold with one tap and one output row: db[k] = sum over blocks of part[block][k]
  return wgrad_fold(s, (float *)workspace, db, blocks, 1, c, 1, c);
}

int lbm3d_distance_periodic(void *stream, const int8_t *solid, float *dist, float *work, int32_t n, int32_t d,
                            int32_t h, int32_t w, int32_t margin) {
  if (n < 1 || d < 1 || h < 1 || w < 1) return cfail(LBM3D_EINVAL, "empty volume");
  if (margin < 1 || margin > 1024) return cfail(LBM3D_EINVAL, "margin must be 1..1024");
  if (!solid || !dist || !work) return cfail(LBM3D_EINVAL, "NULL tensor");
  const int64_t total = (int64_t)n * d * h * w;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  // innermost axis from the solid mask into dist, then the middle axis into
  // work and the outer axis back into dist (a pass cannot run in place: it
  // reads values other threads overwrite)
  k_dist_pass<<<blocks, 256, 0, s>>>(solid, nullptr, dist, total, d, h, w, 2, margin, 0);
  k_dist_pass<<<blocks, 256, 0, s>>>(nullptr, dist, work, total, d, h, w, 1, margin, 0);
  k_dist_pass<<<blocks, 256, 0, s>>>(nullptr, work, dist, total, d, h, w, 0, margin, 1);
  return ccheck("lbm3d_distance_periodic");
}

}  // extern "C"
