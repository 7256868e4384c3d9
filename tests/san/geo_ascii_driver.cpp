// Sanitizer driver for the library's ASCII geometry reader
// (taichi_lbm3d_amd/csrc/geo_ascii.h, behind lbm3d_read_geo_ascii): feeds it
// well-formed and malformed files and checks each result.  Built with ASan +
// UBSan by tests/san/Makefile; a report or a wrong result exits non-zero.
// Usage: geo_ascii_san <scratch directory>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "geo_ascii.h"

static int failures = 0;

static int64_t parse(const std::string &dir, const char *name, const std::string &text, int nx, int ny, int nz,
                     std::vector<int8_t> &out) {
  const std::string path = dir + "/" + name;
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return -100;
  fwrite(text.data(), 1, text.size(), f);
  fclose(f);
  out.assign((size_t)nx * ny * nz, 99);
  char err[512] = "";
  const int64_t rc = lbm3d_ascii::read_geo(path.c_str(), nx, ny, nz, out.data(), err, sizeof(err));
  printf("%-16s rc %lld %s\n", name, (long long)rc, err);
  return rc;
}

static void expect(bool ok, const char *what) {
  if (!ok) {
    printf("FAILED: %s\n", what);
    ++failures;
  }
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  std::vector<int8_t> s;
  // x fastest: value n is cell (n % 2, (n / 2) % 3, n / 6) of a 2 x 3 x 2 grid
  std::string t;
  for (int n = 0; n < 12; ++n) t += std::to_string(n % 3 == 0 ? 1 : 0) + "\n";
  expect(parse(dir, "plain", t, 2, 3, 2, s) == 12, "plain count");
  for (int n = 0; n < 12; ++n) {
    const int i = n % 2, j = (n / 2) % 3, k = n / 6;
    expect(s[(i * 3 + j) * 2 + k] == (n % 3 == 0 ? 1 : 0), "plain layout");
  }
  expect(parse(dir, "separators", "0,1\t0\r\n1 0 0\n1\n0\n0\n1\n0\n0", 2, 3, 2, s) == 12, "separators");
  expect(parse(dir, "no_newline", "0 0 0 0 0 0 0 0 0 0 0 7", 2, 3, 2, s) == 12 && s[(1 * 3 + 2) * 2 + 1] == 1,
         "last value without newline");
  expect(parse(dir, "fractions", "0.5\n-0.5\n2e0\n-1.9\n-300\n-1e300\n+0\n1e-300\n-0.0\n3.25\n0\n0\n", 2, 3, 2, s) ==
             12,
         "fractions");
  expect(s[0] == 1 && s[6] == 0 && s[2] == 1 && s[8] == -1, "fraction values");
  expect(s[4] == (int8_t)(-300 & 0xff) || s[4] == -44, "int8 wrap of -300");
  std::string longtok = "0." + std::string(120, '0') + "1";
  t = longtok;
  for (int n = 1; n < 12; ++n) t += " 0";
  expect(parse(dir, "long_token", t, 2, 3, 2, s) == 12 && s[0] == 1, "a 123-character number");
  expect(parse(dir, "too_few", "0 1 0", 2, 3, 2, s) == LBM3D_EINVAL, "too few values");
  expect(parse(dir, "too_many", std::string(13 * 2, '0').replace(1, std::string::npos, " 0 0 0 0 0 0 0 0 0 0 0 0 0"),
               2, 3, 2, s) == LBM3D_EINVAL,
         "too many values");
  expect(parse(dir, "garbage", "0 1 abc 0 0 0 0 0 0 0 0 0", 2, 3, 2, s) == LBM3D_EINVAL, "not a number");
  expect(parse(dir, "suffix", "0 1 1.5abc 0 0 0 0 0 0 0 0 0", 2, 3, 2, s) == LBM3D_EINVAL, "number with a suffix");
  expect(parse(dir, "neg_inf", "0 1 -inf 0 0 0 0 0 0 0 0 0", 2, 3, 2, s) == LBM3D_EINVAL, "-inf");
  expect(parse(dir, "nan", "0 1 nan 0 0 0 0 0 0 0 0 0", 2, 3, 2, s) == LBM3D_EINVAL, "nan");
  expect(parse(dir, "empty", "", 1, 1, 1, s) == LBM3D_EINVAL, "empty file");
  expect(parse(dir, "blank", " \n\n\t ", 1, 1, 1, s) == LBM3D_EINVAL, "whitespace only");
  expect(parse(dir, "sign_only", "-", 1, 1, 1, s) == LBM3D_EINVAL, "a lone sign");
  expect(parse(dir, "long_digits", "12345678901234567890123", 1, 1, 1, s) == 1 && s[0] == 1, "23 digits");
  char err[512] = "";
  std::vector<int8_t> one(1);
  expect(lbm3d_ascii::read_geo(dir.c_str(), 1, 1, 1, one.data(), err, sizeof(err)) == LBM3D_EINVAL, "a directory");
  expect(lbm3d_ascii::read_geo((dir + "/missing").c_str(), 1, 1, 1, one.data(), err, sizeof(err)) == LBM3D_EINVAL,
         "a missing file");
  expect(lbm3d_ascii::read_geo(nullptr, 1, 1, 1, one.data(), err, sizeof(err)) == LBM3D_EINVAL, "NULL path");
  expect(lbm3d_ascii::read_geo((dir + "/plain").c_str(), 0, 1, 1, one.data(), err, sizeof(err)) == LBM3D_EINVAL,
         "zero extent");
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
