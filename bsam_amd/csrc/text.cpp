// text.cpp -- the number formats of the frame files (WriteQ2D/WriteQ3D,
// lib/bsaminputoutput.f90:138-168): rows of values written es20.12 (2D) or
// f25.12 (3D), two blanks apart, one row a line.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "bsam_gpu.h"

namespace {

// Fortran ES20.12: d.dddddddddddd followed by E+nn, or by +nnn without the
// letter E once the exponent needs three digits; right-aligned in 20.
int es20(double v, char *out) {
  char b[64];
  int n = snprintf(b, sizeof b, "%.12E", v);
  char *e = strchr(b, 'E');
  if (e && (strlen(e) > 4)) {   // E+ddd: drop the letter
    memmove(e, e + 1, strlen(e));
    n -= 1;
  }
  if (n > 20) n = 20;
  int pad = 20 - n;
  memset(out, ' ', pad);
  memcpy(out + pad, b, n);
  return 20;
}

int f25(double v, char *out) {
  char b[64];
  int n = snprintf(b, sizeof b, "%25.12f", v);
  if (n > 25) {   // Fortran fills a field too narrow with asterisks
    memset(out, '*', 25);
    return 25;
  }
  memcpy(out, b, 25);
  return 25;
}

}  // namespace

extern "C" long bsam_format_rows(const double *vals, long nrows, int ncols, long rowstride,
                                 long colstride, int kind, char *out, long cap) {
  if (!vals || !out || nrows < 0 || ncols < 1) return -1;
  const int w = kind == 0 ? 20 : 25;
  const long line = (long)ncols * (w + 2) - 2 + 1;   // every row has this width
  const long need = nrows * line;
  if (cap < need) return -1;
  auto rows = [&](long r0, long r1) {
    for (long r = r0; r < r1; r++) {
      char *p = out + r * line;
      for (int c = 0; c < ncols; c++) {
        if (c) {
          *p++ = ' ';
          *p++ = ' ';
        }
        const double v = vals[r * rowstride + c * colstride];
        p += kind == 0 ? es20(v, p) : f25(v, p);
      }
      *p = '\n';
    }
  };
  // rows of fixed width: split them over host threads
  const long nt = std::max<long>(1, std::min<long>(std::min<long>(16, (long)std::thread::hardware_concurrency()),
                                                   nrows / 4096));
  if (nt <= 1) {
    rows(0, nrows);
  } else {
    std::vector<std::thread> th;
    for (long t = 0; t < nt; t++)
      th.emplace_back(rows, nrows * t / nt, nrows * (t + 1) / nt);
    for (auto &x : th) x.join();
  }
  return need;
}
