/*
 * fixsum_host.cpp -- runs the kernel set's exact fixed-point accumulator
 * (FixSum and fixed_total from reduce_common.h, compiled for the host) on
 * values read from stdin as C99 hex floats, one per line, and prints the
 * 192-bit accumulator words and the converted double. The words are split
 * into 32-bit limbs and recombined as the march kernels' epilogue does.
 */
#include <cstdio>
#include <cstdlib>

#include "../../flow_amd/csrc/kernels/reduce_common.h"

int main() {
  flowhip::FixSum f{0, 0, 0, 0};
  char line[128];
  while (std::fgets(line, sizeof(line), stdin)) f.add(std::strtod(line, nullptr));
  const unsigned long long t[7] = {f.w0 & 0xffffffffull, f.w0 >> 32,
                                   f.w1 & 0xffffffffull, f.w1 >> 32,
                                   f.w2 & 0xffffffffull, f.w2 >> 32,
                                   (unsigned long long)f.bad};
  std::printf("%llu %llu %llu %u %a\n", f.w0, f.w1, f.w2, f.bad,
              flowhip::fixed_total(t));
  return 0;
}
