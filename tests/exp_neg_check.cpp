// Host-side accuracy check of exp_neg (fml_amd/csrc/exp_neg.h, the exp of
// the ARAD overlap) against long-double expl on (-16, 0], and of
// exp_neg_wide on [-760, 1e-9].  Prints two lines: points, max ulp error,
// worst x.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

#define FMLH_HD
#include "exp_neg.h"

static const double kTab[128] = {FMLH_EXP2_128_VALUES};

static double ulp_err(double v, long double ref) {
  const double rd = (double)ref;
  const double u = std::nextafter(std::fabs(rd), INFINITY) - std::fabs(rd);
  return (double)(std::fabs((long double)v - ref) / (long double)u);
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  double worst = 0.0, wx = 0.0;
  long count = 0;
  auto check = [&](double x) {
    if (!(x > -16.0) || x > 0.0) return;
    const double v = fmlh::exp_neg(x, kTab);
    const double e = ulp_err(v, expl((long double)x));
    ++count;
    if (e > worst) {
      worst = e;
      wx = x;
    }
  };
  // a uniform grid, and the neighbourhood of every reduction boundary
  for (long i = 0; i < n; ++i) check(-16.0 * ((double)i + 0.5) / (double)n);
  const double step = 0.6931471805599453 / 128.0;
  for (int k = 0; k >= -2960; --k) {
    const double c = (k + 0.5) * step;
    double x = c;
    for (int s = 0; s < 8; ++s) {
      check(x);
      x = std::nextafter(x, 1.0);
    }
    x = c;
    for (int s = 0; s < 8; ++s) {
      x = std::nextafter(x, -1.0);
      check(x);
    }
    check(k * step);
  }
  check(-0.0);
  check(0.0);
  check(std::nextafter(-16.0, 0.0));
  std::printf("%ld %.6f %.17g\n", count, worst, wx);
  // exp_neg_wide on [-760, 1e-9]: grid, both sides of the -707 switch, the
  // subnormal range and tiny positive arguments
  double worst_w = 0.0, wxw = 0.0;
  long count_w = 0;
  auto check_w = [&](double x) {
    const double v = fmlh::exp_neg_wide(x, kTab);
    const long double ref = expl((long double)x);
    double e;
    if ((double)ref == 0.0)
      e = v == 0.0 ? 0.0 : 1e9;
    else
      e = ulp_err(v, ref);
    ++count_w;
    if (e > worst_w) {
      worst_w = e;
      wxw = x;
    }
  };
  for (long i = 0; i < n; ++i) check_w(-760.0 * ((double)i + 0.5) / (double)n);
  for (double c : {-707.0, -708.0, -708.3964185322641, -745.1332191019411}) {
    double x = c;
    for (int s = 0; s < 64; ++s) {
      check_w(x);
      x = std::nextafter(x, 1.0);
    }
    x = c;
    for (int s = 0; s < 64; ++s) {
      x = std::nextafter(x, -1e300);
      check_w(x);
    }
  }
  for (double x = 1e-300; x < 1e-9; x *= 3.0) check_w(x);
  std::printf("%ld %.6f %.17g\n", count_w, worst_w, wxw);
  return 0;
}
