// Host build of csrc/k_f64math.h against the C library's FP64 log and pow:
// every FP32 input of a range, both results rounded to FP32 and compared;
// also the largest FP64 difference in ulp.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "k_f64math.h"

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static long long ulp_diff(double a, double b) {
  int64_t x, y;
  memcpy(&x, &a, 8);
  memcpy(&y, &b, 8);
  return (long long)(x > y ? x - y : y - x);
}

int main() {
  {
    long n = 0, mis = 0;
    long long worst = 0;
    for (uint32_t u = to_bits(0.5f); u < to_bits(2.0f); u++) {
      const float x = from_bits(u);
      const double a = f64_log((double)x), b = log((double)x);
      mis += (float)a != (float)b;
      const long long d = ulp_diff(a, b);
      worst = d > worst ? d : worst;
      n++;
    }
    printf("log checked=%ld mismatches=%ld max_ulp=%lld\n", n, mis, worst);
  }
  {
    const double c = (double)1.4f;
    long n = 0, mis = 0;
    long long worst = 0;
    for (uint32_t u = to_bits(0.0625f); u < to_bits(2.0f); u++) {
      const float t = from_bits(u);
      const double a = f64_pow((double)t, c), b = pow((double)t, c);
      mis += (float)a != (float)b;
      const long long d = ulp_diff(a, b);
      worst = d > worst ? d : worst;
      n++;
    }
    printf("pow checked=%ld mismatches=%ld max_ulp=%lld\n", n, mis, worst);
  }
  return 0;
}
