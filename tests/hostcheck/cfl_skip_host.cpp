/*
 * cfl_skip_host.cpp -- cfl_above from the kernel set's hydro_math.h, compiled
 * for the host, against the CFL candidates it lets the Lagrangian marches
 * leave out (cell_timestep, FP64; cell_timestep_sq, the FP32-state form).
 * Whenever cfl_above says a cell cannot lower a running minimum L, the
 * candidate must be defined (not NaN) and lesser(L, candidate) must be L bit
 * for bit. Inputs: random cells over many decades, limits set within a few
 * ulps of each cell's own candidate on both sides, and special values
 * (zeros, negatives, infinities, NaN, subnormals, extremes). Prints the
 * cases checked, the cases cfl_above passed, and the violations (must be 0)
 * for each form.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>

#include "../../flow_amd/csrc/kernels/hydro_math.h"

using flowhip::cell_timestep;
using flowhip::cell_timestep_sq;
using flowhip::cfl_above;
using flowhip::lesser;

static uint64_t state = 0x9e3779b97f4a7c15ull;
static uint64_t next() {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return state;
}
static double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
static double logu(double lo, double hi) { return std::exp(std::log(lo) + uni() * std::log(hi / lo)); }

static long n64 = 0, pass64 = 0, bad64 = 0, n32 = 0, pass32 = 0, bad32 = 0;

static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

static void check(double e, double qx, double qy, double rho, double cdx, double cdy,
                  double L, double Lsq) {
  ++n64;
  if (cfl_above(e, qx, qy, rho, cdx, cdy, L * L)) {
    ++pass64;
    const double c = cell_timestep(e, qx, qy, rho, cdx, cdy);
    if (!(c == c) || !same(lesser(L, c), L)) ++bad64;
  }
  ++n32;
  if (cfl_above(e, qx, qy, rho, cdx, cdy, Lsq)) {
    ++pass32;
    const double c = cell_timestep_sq(e, qx, qy, rho, cdx, cdy);
    if (!(c == c) || !same(lesser(Lsq, c), Lsq)) ++bad32;
  }
}

int main() {
  const double inf = std::numeric_limits<double>::infinity();
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const double tiny = std::numeric_limits<double>::denorm_min();
  /* random cells and limits near their own candidates */
  for (int k = 0; k < 2000000; ++k) {
    const double rho = logu(1e-12, 1e6);
    const double e = logu(1e-10, 1e8);
    const double qx = uni() < 0.3 ? 0.0 : logu(1e-14, 1e8);
    const double qy = uni() < 0.3 ? 0.0 : logu(1e-14, 1e8);
    const double cdx = logu(1e-7, 1e1), cdy = uni() < 0.5 ? cdx : logu(1e-7, 1e1);
    const double c = cell_timestep(e, qx, qy, rho, cdx, cdy);
    const double cs = cell_timestep_sq(e, qx, qy, rho, cdx, cdy);
    for (int j = -8; j <= 8; ++j) {
      /* L and Lsq within j ulps-scaled steps of the candidates, and far off */
      const double f = 1.0 + j * 1.1e-16 * (1 + (k % 5) * 1000);
      check(e, qx, qy, rho, cdx, cdy, c * f, cs * f);
    }
    /* limits right at the test's margin (1e-12 on the squares, 5e-13 on dt) */
    for (int j = 0; j <= 10; ++j) {
      const double t = 0.9 + 0.02 * j;
      check(e, qx, qy, rho, cdx, cdy, c * (1.0 - 5.0e-13 * t), cs * (1.0 - 1.0e-12 * t));
    }
    check(e, qx, qy, rho, cdx, cdy, c * 0.5, cs * 0.25);
    check(e, qx, qy, rho, cdx, cdy, std::nextafter(c, 0.0), std::nextafter(cs, 0.0));
    check(e, qx, qy, rho, cdx, cdy, std::nextafter(c, inf), std::nextafter(cs, inf));
  }
  /* special values in every argument */
  const double sp[] = {0.0, -0.0, 1.0, -1.0, 1e-300, -1e-300, 1e300, -1e300, tiny, -tiny,
                       inf, -inf, nan, 2.5, 1e-250, 1e250, 1e-280, 1e280};
  const int ns = sizeof sp / sizeof sp[0];
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b)
      for (int d = 0; d < ns; ++d)
        for (int g = 0; g < 6; ++g) {
          const double base[6] = {1.0, 0.1, 0.2, 1.0, 1e-3, 1e-3};
          double v[6] = {base[0], base[1], base[2], base[3], base[4], base[5]};
          v[g] = sp[a];
          v[(g + 1) % 6] = sp[b];
          const double L = sp[d];
          check(v[0], v[1], v[2], v[3], v[4], v[5], L, L);
          check(v[0], v[1], v[2], v[3], v[4], v[5], std::fabs(L) * 1e-6, std::fabs(L) * 1e-12);
        }
  std::printf("%ld %ld %ld %ld %ld %ld\n", n64, pass64, bad64, n32, pass32, bad32);
  return 0;
}
