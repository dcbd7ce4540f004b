// Text formatting of the snapshot and ic.csv lines (io.cpp:59-89: "%d,%d,%d,%g,%g,%g,%g,%d\n"), without printf.
// std::to_chars with chars_format::general and precision 6 is specified to print exactly what printf("%g") prints
// in the C locale, and integers as "%d"; it needs no locale or stream state, so threads can format disjoint ranges
// of a lattice at once.
#ifndef MCLICS_CSVFMT_H_
#define MCLICS_CSVFMT_H_

#include <charconv>
#include <cstddef>

namespace mclcsv {

// Longest line: 4 ints of at most 11 characters, 4 "%g" values of at most 13, 7 commas and a newline.
constexpr size_t kMaxLine = 4 * 11 + 4 * 13 + 8;

inline char *put_int(char *p, int v) { return std::to_chars(p, p + 11, v).ptr; }

// printf("%g", (double)v): a float argument of printf is promoted to double.
inline char *put_g(char *p, float v) {
  return std::to_chars(p, p + 13, (double)v, std::chars_format::general, 6).ptr;
}

// One line "i,j,k,nx,ny,nz,S,pt\n"; returns the end of what was written (at most kMaxLine characters).
inline char *put_line(char *p, int i, int j, int k, float nx, float ny, float nz, float S, int pt) {
  p = put_int(p, i);
  *p++ = ',';
  p = put_int(p, j);
  *p++ = ',';
  p = put_int(p, k);
  *p++ = ',';
  p = put_g(p, nx);
  *p++ = ',';
  p = put_g(p, ny);
  *p++ = ',';
  p = put_g(p, nz);
  *p++ = ',';
  p = put_g(p, S);
  *p++ = ',';
  p = put_int(p, pt);
  *p++ = '\n';
  return p;
}

}  // namespace mclcsv

#endif
