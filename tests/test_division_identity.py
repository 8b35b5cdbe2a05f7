"""The multigrid sweeps divide by each cell's diagonal with Markstein's
correction step from a precomputed reciprocal (div_rcp in
campfire_amd/csrc/cf_device.h): y = RN(1/d), q = RN(x*y),
r = fma(-q, d, x), result fma(r, y, q).  The kernels' bitwise parity with
the oracle's true divisions rests on this being the correctly rounded x/d.
Checked here in C (gcc, libm's fmaf) on every significand of x against
divisors with long runs of ones and zeros, on random pairs over the exponent
range the sweeps see, and on dividends down through the subnormal range (the
IEEE-division branch) and signed zeros."""
import os
import subprocess
import tempfile

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
/* div_rcp of campfire_amd/csrc/cf_device.h */
static float div_rcp(float x, float d, float y) {
  const float q = x * y;
  const float m = copysignf(fmaf(fmaf(-q, d, x), y, q), x);
  uint32_t ax;
  memcpy(&ax, &x, 4);
  ax &= 0x7fffffffu;
  if (ax - 1u < 0x0d7fffffu) return x / d;
  return m;
}
static int check(float x, float d) {
  volatile float yv = 1.0f / d;
  const float v = div_rcp(x, d, yv), ref = x / d;
  return memcmp(&v, &ref, 4) != 0;
}
int main(void) {
  long bad = 0;
  uint32_t divs[48];
  for (int i = 0; i < 12; i++) divs[i] = 0x7fffffu >> i;
  for (int i = 0; i < 12; i++) divs[12 + i] = (0x7fffffu << (i + 1)) & 0x7fffffu;
  for (int i = 0; i < 24; i++) divs[24 + i] = (uint32_t)(i * 349525u + 12345u) & 0x7fffffu;
  for (int b = 0; b < 48; b++) {
    const float d = bits((127u << 23) | divs[b]);
    for (uint32_t m = 0; m <= 0x7fffffu; m += 3) bad += check(bits((127u << 23) | m), d);
  }
  uint64_t s = 88172645463325252ull;
  for (long n = 0; n < 4000000; n++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    const uint32_t ex = 127u - 40u + (uint32_t)(s % 81), ed = 127u - 40u + (uint32_t)((s >> 8) % 81);
    const float x = bits((ex << 23) | ((uint32_t)(s >> 16) & 0x7fffffu)) * ((s >> 63) ? -1.f : 1.f);
    const float d = bits((ed << 23) | ((uint32_t)(s >> 40) & 0x7fffffu));
    bad += check(x, d);
  }
  /* dividends down to the subnormal range and signed zeros, divisors 2^-20..2^20 */
  for (long n = 0; n < 4000000; n++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    const uint32_t ex = (uint32_t)(s % 40), ed = 127u - 20u + (uint32_t)((s >> 8) % 41);
    const float x = bits((ex << 23) | ((uint32_t)(s >> 16) & 0x7fffffu)) * ((s >> 63) ? -1.f : 1.f);
    const float d = bits((ed << 23) | ((uint32_t)(s >> 40) & 0x7fffffu));
    bad += check(x, d);
  }
  bad += check(0.0f, 3.0f) + check(-0.0f, 3.0f) + check(0x1p-149f, 1.5f) + check(-0x1p-127f, 0.7f);
  printf("%ld\n", bad);
  return 0;
}
"""


def test_markstein_division_is_correctly_rounded():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "div.c"), os.path.join(d, "div")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0
