/* Host restatement of the device division div_const (wrfdc_common.h), checked
 * bitwise against the CPU's IEEE division x / c for c = 12 and c = 60.
 *
 * Inputs: every float with biased exponent below 17 (every subnormal and the
 * smallest normals, which hold the rare path and its boundary), both signs,
 * plus every `stride`-th bit pattern of the rest (default 251).  Built with gcc -O2 -mfma
 * -ffp-contract=off; the correction term t is kept in a volatile so the
 * compiler cannot rewrite -(q*c - x) as x - q*c (that changes the sign of a
 * zero result, which the device code keeps).
 *
 * usage: div_const_check [stride]   prints "c=<c> checked=<n> mismatches=<m>"
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bits_to_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t float_to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float div_tiny(float xv, float c, float rc) {
  const float X = ldexpf(ldexpf(ldexpf(fabsf(xv), 64), 64), 21);
  float n = rintf(X * rc);
  float r = fmaf(-n, c, X);
  n = n + rintf(r * rc);
  r = fmaf(-n, c, X);
  const float h = 0.5f * c;
  const int odd = (n - 2.f * rintf(0.5f * n)) != 0.f;
  if (r > h || (r == h && odd)) n = n + 1.f;
  else if (r < -h || (r == -h && odd)) n = n - 1.f;
  return copysignf(n * 0x1p-149f, xv);
}

static float div_const(float x, float c, float rc) {
  const float q = x * rc;
  volatile float t = fmaf(q, c, -x);
  float q2 = fmaf(-t, rc, q);
  if (fpclassify(q2) == FP_SUBNORMAL) q2 = div_tiny(x, c, rc);
  return q2;
}

int main(int argc, char **argv) {
  const long stride = argc > 1 ? atol(argv[1]) : 251;
  const float cs[2] = {12.f, 60.f};
  const long n_low = 2L * 17L << 23; /* both signs, biased exponents 0..16 */
  const long n_rest = (1L << 32) / stride + 1;
  int fail = 0;
  for (int ci = 0; ci < 2; ++ci) {
    const float c = cs[ci], rc = 1.f / c;
    long checked = 0, bad = 0;
#pragma omp parallel for reduction(+ : checked, bad) schedule(static)
    for (long t = 0; t < n_low + n_rest; ++t) {
      uint32_t u;
      if (t < n_low) {
        u = (uint32_t)((t >> 23) / 17) << 31 | (uint32_t)((t >> 23) % 17) << 23 | (uint32_t)(t & 0x7fffff);
      } else {
        const long b = (t - n_low) * stride;
        if (b >= (1L << 32) || ((b >> 23) & 0xff) < 17) continue;
        u = (uint32_t)b;
      }
      const float x = bits_to_float(u);
      if (!isfinite(x)) continue;
      ++checked;
      if (float_to_bits(div_const(x, c, rc)) != float_to_bits(x / c)) ++bad;
    }
    printf("c=%g checked=%ld mismatches=%ld\n", c, checked, bad);
    fail |= bad != 0;
  }
  return fail;
}
