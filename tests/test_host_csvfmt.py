"""The mclics program's snapshot line formatter (mclics_amd/host/csvfmt.h) against printf.

The snapshot and ic.csv writer (io.cpp:59-89 prints "%d,%d,%d,%g,%g,%g,%g,%d\\n") formats its lines with
std::to_chars on several threads. This test compiles a small program against the header and compares every line
with snprintf of the same format, on random directors, local order values of every magnitude, NaN (empty sites),
infinities, signed zeros, subnormals and extreme integers. Byte-identical output is required.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include "csvfmt.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>

int main() {
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  const float special[] = {0.0f, -0.0f, 1.0f, -1.0f, NAN, -NAN, INFINITY, -INFINITY, 1e-5f, 1e-4f, 99999.95f,
                           999999.5f, 1234567.0f, std::numeric_limits<float>::denorm_min(),
                           std::numeric_limits<float>::min(), std::numeric_limits<float>::max(), 9.9999995e-5f,
                           0.000123456f, 0.5f, 1.0f / 3.0f};
  const int ints[] = {0, 1, -1, 9, 10, 255, 65535, 2147483647, -2147483647 - 1};
  const int ns = sizeof(special) / sizeof(special[0]), ni = sizeof(ints) / sizeof(ints[0]);
  char a[256], b[256];
  long bad = 0, n = 0;
  for (long t = 0; t < 3000000; t++) {
    float v[4];
    for (int c = 0; c < 4; c++) {
      const unsigned pick = (unsigned)(g() % 8);
      if (pick == 0) v[c] = special[g() % ns];
      else if (pick == 1) v[c] = std::ldexp(u(g), (int)(g() % 80) - 60);
      else v[c] = u(g);
    }
    int iv[4];
    for (int c = 0; c < 4; c++) iv[c] = (g() % 4 == 0) ? ints[g() % ni] : (int)(g() % 2048);
    snprintf(a, sizeof a, "%d,%d,%d,%g,%g,%g,%g,%d\n", iv[0], iv[1], iv[2], v[0], v[1], v[2], v[3], iv[3]);
    char *e = mclcsv::put_line(b, iv[0], iv[1], iv[2], v[0], v[1], v[2], v[3], iv[3]);
    if ((size_t)(e - b) > mclcsv::kMaxLine) { printf("line too long\n"); return 2; }
    *e = 0;
    n++;
    if (strcmp(a, b) != 0) {
      if (bad < 5) printf("printf: %s  csvfmt: %s", a, b);
      bad++;
    }
  }
  printf("lines %ld mismatches %ld\n", n, bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_csvfmt_lines_equal_printf(tmp_path):
    src = tmp_path / "csvfmt_check.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "csvfmt_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mclics_amd", "host"), "-o", str(exe),
                    str(src)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
