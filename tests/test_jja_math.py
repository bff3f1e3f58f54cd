"""The step kernel's double-precision sine (pyjjasim_amd/csrc/jja_math.h) against
the C library's sin, compiled for the host with gcc from the same header.

Bounds checked: at most 1 ulp for |x| <= 1e9 on uniform samples; near the zeros
k*pi/2 (where the result is tiny and relative error is meaningless) an absolute
error below 1e-24; exact zeros and NaN propagation."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "pyjjasim_amd", "csrc", "jja_math.h")

PROG = r"""
#include "%s"
#include <stdio.h>
#include <stdint.h>
#include <string.h>
static int64_t ulps(double a, double b) {
  if (a == b) return 0;
  int64_t ia, ib; memcpy(&ia, &a, 8); memcpy(&ib, &b, 8);
  if (ia < 0) ia = INT64_MIN - ia;
  if (ib < 0) ib = INT64_MIN - ib;
  int64_t d = ia - ib; return d < 0 ? -d : d;
}
int main(void) {
  const double ranges[] = {1e-8, 1.0, 4.0, 100.0, 1e4, 1e6, 1e9};
  unsigned long long s = 88172645463325252ULL;
  int64_t worst = 0;
  for (int ri = 0; ri < 7; ++ri)
    for (long i = 0; i < 400000; ++i) {
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      double x = (2.0 * ((double)(s >> 11) * 0x1p-53) - 1.0) * ranges[ri];
      int64_t d = ulps(jja_sin(x), sin(x));
      if (d > worst) worst = d;
    }
  double abserr = 0.0;
  for (long k = 1; k < 300000; ++k) {
    double x = k * 1.5707963267948966;
    for (int d = -2; d <= 2; ++d) {
      double y = x;
      for (int e = 0; e < (d < 0 ? -d : d); ++e) y = nextafter(y, d < 0 ? 0.0 : 1e300);
      double a = fabs(jja_sin(y) - sin(y));
      if (a > abserr) abserr = a;
    }
  }
  int specials = jja_sin(0.0) == 0.0 && jja_sin(NAN) != jja_sin(NAN) && jja_sin(INFINITY) != jja_sin(INFINITY);
  int64_t worst_sc = 0;
  for (long i = 0; i < 400000; ++i) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    double x = (2.0 * ((double)(s >> 11) * 0x1p-53) - 1.0) * 1e4, sn, cs;
    jja_sincos(x, &sn, &cs);
    int64_t d1 = ulps(sn, sin(x)), d2 = fabs(cos(x)) > 1e-3 ? ulps(cs, cos(x)) : 0;
    if (fabs(cs - cos(x)) > 1e-15) d2 = 1000;
    if (d1 > worst_sc) worst_sc = d1;
    if (d2 > worst_sc) worst_sc = d2;
  }
  printf("%%lld %%.6g %%d %%lld\n", (long long)worst, abserr, specials, (long long)worst_sc);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_kernel_sine_matches_libm(tmp_path):
    src = tmp_path / "t.c"
    src.write_text(PROG % HEADER)
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    worst_ulp, abserr, specials, worst_sc = int(out[0]), float(out[1]), int(out[2]), int(out[3])
    assert worst_ulp <= 1, f"max error {worst_ulp} ulp"
    assert worst_sc <= 1, f"sincos: max error {worst_sc} ulp"
    assert abserr < 1e-24, f"absolute error near k*pi/2: {abserr}"
    assert specials == 1
