// Host-only parser of the reference's ASCII voxel file (init_geo, SP:173-177):
// whitespace-separated numbers, x fastest, then y, then z.  Plain C++ with no
// HIP dependency, so the library (geo_io.hip) and a sanitizer build of a
// stand-alone driver compile the same code.
#ifndef LBM3D_GEO_ASCII_H
#define LBM3D_GEO_ASCII_H

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <sys/stat.h>

#include "lbm3d.h"

namespace lbm3d_ascii {

inline bool is_sep(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == ','; }

// next number of a whitespace-separated ASCII stream; false at end of input.
// A token that is not a number yields NaN.
inline bool next_number(const char *&p, const char *end, double &val) {
  while (p < end && is_sep(*p)) ++p;
  if (p >= end) return false;
  // fast path: an optionally signed integer of at most 18 digits followed by a separator
  const char *q = p;
  bool neg = false;
  if (*q == '-' || *q == '+') neg = *q++ == '-';
  int64_t acc = 0;
  const char *digits = q;
  while (q < end && *q >= '0' && *q <= '9' && q - digits < 18) acc = acc * 10 + (*q++ - '0');
  if (q > digits && (q == end || is_sep(*q))) {
    val = neg ? -(double)acc : (double)acc;
    p = q;
    return true;
  }
  // general number (fractions, exponents): strtod on a terminated copy of the token
  size_t len = 0;
  while (p + len < end && !is_sep(p[len])) ++len;
  char small[64];
  char *buf = len < sizeof(small) ? small : (char *)malloc(len + 1);
  if (!buf) {
    val = NAN;
    p += len;
    return true;
  }
  memcpy(buf, p, len);
  buf[len] = 0;
  char *stop = nullptr;
  errno = 0;
  val = strtod(buf, &stop);
  if (stop == buf || (size_t)(stop - buf) != len) val = NAN;  // not (entirely) a number
  if (buf != small) free(buf);
  p += len;
  return true;
}

// Parse `path` into the host int8 array solid[nx][ny][nz] (C order).  Values
// > 0 become 1, the others truncate toward zero to int8 (np.loadtxt then
// astype, SP:175-177).  Returns the number of values, or a negative LBM3D_E*
// code with a message in err.
inline int64_t read_geo(const char *path, int32_t nx, int32_t ny, int32_t nz, int8_t *solid, char *err,
                        size_t err_len) {
  if (!path || !solid || nx < 1 || ny < 1 || nz < 1) {
    snprintf(err, err_len, "bad argument to lbm3d_read_geo_ascii");
    return LBM3D_EINVAL;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    snprintf(err, err_len, "cannot open %s", path);
    return LBM3D_EINVAL;
  }
  struct stat st;
  if (fstat(fileno(f), &st) != 0 || !S_ISREG(st.st_mode)) {  // directories, pipes, devices
    fclose(f);
    snprintf(err, err_len, "not a regular file: %s", path);
    return LBM3D_EINVAL;
  }
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    snprintf(err, err_len, "cannot seek %s", path);
    return LBM3D_EINVAL;
  }
  const long size = ftell(f);
  if (size < 0) {  // not a seekable regular file
    fclose(f);
    snprintf(err, err_len, "cannot size %s", path);
    return LBM3D_EINVAL;
  }
  rewind(f);
  const size_t cap = size > 0 ? (size_t)size : 1;
  char *buf = (char *)malloc(cap);
  if (!buf) {
    fclose(f);
    snprintf(err, err_len, "out of host memory reading %s", path);
    return LBM3D_ESIZE;
  }
  const size_t got = fread(buf, 1, (size_t)size, f);  // size <= cap
  fclose(f);
  const int64_t total = (int64_t)nx * ny * nz;
  const char *p = buf, *end = buf + got;
  int64_t count = 0;
  double v;
  while (next_number(p, end, v)) {
    if (std::isnan(v) || (std::isinf(v) && v < 0)) {  // no int8 value to give these
      free(buf);
      snprintf(err, err_len, "not a number (or -inf) in %s", path);
      return LBM3D_EINVAL;
    }
    if (count < total) {
      // value n of the file is cell (i, j, k) with n = i + nx*(j + ny*k) (x fastest)
      const int64_t i = count % nx, t = count / nx;
      const int64_t j = t % ny, k = t / ny;
      // > 0 -> 1; otherwise truncate toward zero and keep the low 8 bits, as
      // an integer cast to int8 does (fmod first, so huge values stay defined)
      const double c = v > 0 ? 1.0 : std::fmod(std::trunc(v), 256.0);
      solid[(i * ny + j) * nz + k] = (int8_t)(int)c;
    }
    ++count;
  }
  free(buf);
  if (count != total) {
    snprintf(err, err_len, "%s: %lld values, expected %lld", path, (long long)count, (long long)total);
    return LBM3D_EINVAL;
  }
  return count;
}

}  // namespace lbm3d_ascii

#endif  // LBM3D_GEO_ASCII_H
