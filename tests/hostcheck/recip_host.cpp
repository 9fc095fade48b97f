/*
 * recip_host.cpp -- runs recip_pow2 from the kernel set's common.h, compiled
 * for the host, on every power of two the power-of-two mesh path admits
 * (2^k, |k| <= 300 in FP64, |k| <= 40 in FP32) and compares its bits with
 * the IEEE quotient 1/x. Prints, per type, the number of inputs checked and
 * the number whose bits differ.
 */
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../flow_amd/csrc/kernels/common.h"

int main() {
  int n64 = 0, bad64 = 0, n32 = 0, bad32 = 0;
  for (int k = -300; k <= 300; ++k) {
    const double x = std::ldexp(1.0, k);
    const volatile double one = 1.0;
    const double q = one / x, r = flowhip::recip_pow2(x);
    ++n64;
    if (std::memcmp(&q, &r, sizeof q) != 0) ++bad64;
  }
  for (int k = -40; k <= 40; ++k) {
    const float x = std::ldexp(1.0f, k);
    const volatile float one = 1.0f;
    const float q = one / x, r = flowhip::recip_pow2(x);
    ++n32;
    if (std::memcmp(&q, &r, sizeof q) != 0) ++bad32;
  }
  std::printf("%d %d %d %d\n", n64, bad64, n32, bad32);
  return 0;
}
