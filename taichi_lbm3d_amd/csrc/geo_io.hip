// Geometry input: the init_geo ASCII reader and an STL voxelizer.  The C ABI is
// declared in include/lbm3d_io.h.
//
// lbm3d_read_geo_ascii parses the reference's voxel file (one number per line,
// x fastest, SP:173-177) on the host with a hand-written number scanner.
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

// next number of a whitespace-separated ASCII stream; false at end of input
bool next_number(const char *&p, const char *end, double &val) {
  while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t' || *p == ',')) ++p;
  if (p >= end) return false;
  // fast path: an optionally signed integer followed by a separator
  const char *q = p;
  bool neg = false;
  if (*q == '-' || *q == '+') neg = *q++ == '-';
  int64_t acc = 0;
  const char *digits = q;
  while (q < end && *q >= '0' && *q <= '9' && q - digits < 18) acc = acc * 10 + (*q++ - '0');
  if (q > digits && (q == end || *q == ' ' || *q == '\n' || *q == '\r' || *q == '\t' || *q == ',')) {
    val = neg ? -(double)acc : (double)acc;
    p = q;
    return true;
  }
  // general number (fractions, exponents): strtod on a bounded copy
  char buf[64];
  size_t len = 0;
  while (p + len < end && len < sizeof(buf) - 1 && p[len] != ' ' && p[len] != '\n' && p[len] != '\r' &&
         p[len] != '\t' && p[len] != ',')
    ++len;
  memcpy(buf, p, len);
  buf[len] = 0;
  char *stop = nullptr;
  errno = 0;
  val = strtod(buf, &stop);
  if (stop == buf) {
    val = NAN;  // not a number: reported by the caller
  }
  p += len;
  return true;
}

}  // namespace

extern "C" {

const char *lbm3d_io_last_error(void) { return io_err; }

int64_t lbm3d_read_geo_ascii(const char *path, int32_t nx, int32_t ny, int32_t nz, int8_t *solid) {
  if (!path || !solid || nx < 1 || ny < 1 || nz < 1) return iofail(LBM3D_EINVAL, "bad argument to lbm3d_read_geo_ascii");
  FILE *f = fopen(path, "rb");
  if (!f) return iofail(LBM3D_EINVAL, "cannot open ", path);
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return iofail(LBM3D_EINVAL, "cannot seek ", path);
  }
  const long size = ftell(f);
  if (size < 0) {  // not a seekable regular file
    fclose(f);
    return iofail(LBM3D_EINVAL, "cannot size ", path);
  }
  rewind(f);
  const size_t cap = size > 0 ? (size_t)size : 1;
  char *buf = (char *)malloc(cap);
  if (!buf) {
    fclose(f);
    return iofail(LBM3D_ESIZE, "out of host memory reading ", path);
  }
  const size_t got = fread(buf, 1, (size_t)size, f);  // size <= cap
  fclose(f);
  const int64_t total = (int64_t)nx * ny * nz;
  const char *p = buf, *end = buf + got;
  int64_t count = 0;
  double v;
  while (next_number(p, end, v)) {
    if (std::isnan(v)) {
      free(buf);
      return iofail(LBM3D_EINVAL, "not a number in ", path);
    }
    if (count < total) {
      // value n of the file is cell (i, j, k) with n = i + nx*(j + ny*k) (x fastest);
      // stored in C order; values > 0 become 1, others truncate to int8
      const int64_t i = count % nx, t = count / nx;
      const int64_t j = t % ny, k = t / ny;
      const double c = v > 0 ? 1.0 : std::trunc(v);
      solid[(i * ny + j) * nz + k] = (int8_t)(int)c;
    }
    ++count;
  }
  free(buf);
  if (count != total) {
    snprintf(io_err, sizeof(io_err), "%s: %lld values, expected %lld", path, (long long)count, (long long)total);
    return LBM3D_EINVAL;
  }
  return count;
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
