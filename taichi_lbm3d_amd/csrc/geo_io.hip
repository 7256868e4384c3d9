// Geometry input: the init_geo ASCII reader and an STL voxelizer.  The C ABI is
// declared in include/lbm3d_io.h.
//
// lbm3d_read_geo_ascii parses the reference's voxel file (one number per line,
// x fastest, SP:173-177) on the host with a hand-written number scanner
// (geo_ascii.h, plain C++ shared with the sanitizer build).
//
// lbm3d_voxelize flags grid nodes inside a closed triangle mesh, the job of the
// reference's Convert_stl_to_binary.cpp:29-157 (libigl's signed distance with
// the fast winding number sign).  Here every node sums the exact generalized
// winding number of all triangles, w(p) = 1/(4 pi) sum_t Omega_t(p), with the
// solid angle of Van Oosterom and Strackee,
//   tan(Omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|),
// a, b, c the triangle's corners relative to p, in FP64; the node is solid when
// w >= 1/2.  One thread per node; triangles are staged through LDS in tiles.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "geo_ascii.h"
#include "lbm3d.h"
#include "lbm3d_io.h"

namespace {

thread_local char io_err[512] = "";

int iofail(int code, const char *msg, const char *what = "") {
  snprintf(io_err, sizeof(io_err), "%s%s", msg, what);
  return code;
}

constexpr int kTriTile = 256;

__global__ void __launch_bounds__(256)
    k_winding(const float *__restrict__ tris, int32_t n_tris, double x0, double y0, double z0, double dx, int32_t nx,
              int32_t ny, int32_t nz, int8_t *__restrict__ solid) {
  __shared__ float s_tri[kTriTile * 9];
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double px = 0, py = 0, pz = 0;
  if (id < n) {
    const int64_t k = id % nz, t = id / nz;
    const int64_t j = t % ny, i = t / ny;
    px = x0 + dx * (double)i;
    py = y0 + dx * (double)j;
    pz = z0 + dx * (double)k;
  }
  double omega = 0.0;
  for (int32_t base = 0; base < n_tris; base += kTriTile) {
    const int cnt = n_tris - base < kTriTile ? n_tris - base : kTriTile;
    __syncthreads();
    for (int e = threadIdx.x; e < cnt * 9; e += blockDim.x) s_tri[e] = tris[(int64_t)base * 9 + e];
    __syncthreads();
    if (id < n) {
      for (int t = 0; t < cnt; ++t) {
        const float *v = s_tri + 9 * t;
        const double ax = v[0] - px, ay = v[1] - py, az = v[2] - pz;
        const double bx = v[3] - px, by = v[4] - py, bz = v[5] - pz;
        const double cx = v[6] - px, cy = v[7] - py, cz = v[8] - pz;
        const double la = sqrt(ax * ax + ay * ay + az * az);
        const double lb = sqrt(bx * bx + by * by + bz * bz);
        const double lc = sqrt(cx * cx + cy * cy + cz * cz);
        const double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
        const double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (bx * cx + by * cy + bz * cz) * la +
                           (cx * ax + cy * ay + cz * az) * lb;
        omega += 2.0 * atan2(det, den);
      }
    }
  }
  if (id < n) {
    const double w = omega / (4.0 * M_PI);
    solid[id] = w >= 0.5 ? 1 : 0;
  }
}

}  // namespace

extern "C" {

const char *lbm3d_io_last_error(void) { return io_err; }

int64_t lbm3d_read_geo_ascii(const char *path, int32_t nx, int32_t ny, int32_t nz, int8_t *solid) {
  return lbm3d_ascii::read_geo(path, nx, ny, nz, solid, io_err, sizeof(io_err));
}

int lbm3d_voxelize(void *stream, const float *tris, int32_t n_tris, double x0, double y0, double z0, double dx,
                   int32_t nx, int32_t ny, int32_t nz, int8_t *solid) {
  if (nx < 1 || ny < 1 || nz < 1 || n_tris < 0 || !(dx > 0)) return iofail(LBM3D_EINVAL, "bad grid or mesh size");
  if ((n_tris > 0 && !tris) || !solid) return iofail(LBM3D_EINVAL, "NULL mesh or output");
  const int64_t n = (int64_t)nx * ny * nz;
  if (n > INT32_MAX) return iofail(LBM3D_ESIZE, "grid has more than 2^31-1 nodes");
  k_winding<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(tris, n_tris, x0, y0, z0, dx, nx, ny, nz,
                                                                         solid);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return iofail(LBM3D_EHIP, "lbm3d_voxelize: ", hipGetErrorString(e));
  return LBM3D_OK;
}

}  // extern "C"
