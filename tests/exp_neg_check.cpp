// Host-side check of fml_amd/csrc/exp_neg.h against the host C library
// (what the reference's Fortran calls) and long-double expl/sinl:
//   exp_neg  on (-16, 0] (the ARAD overlap's exponents),
//   exp_libm on [-760, 1e-9] (the atom-pair kernels' exponents),
//   sin_libm on [0.126, pi/2] and on the ARAD domains d pi/(2 cut),
//   sin_libm on [0, 0.126) (where sin_cr stands in for glibc's Taylor branch),
//   sin_cr   on [0, pi/2].
// One line per set:
//   name points max_ulp worst_x libm_differs
// max_ulp: the largest error in ulps against the long-double value;
// libm_differs: results that differ from the host libm's.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <random>

#define FMLH_HD
#include "exp_neg.h"

static const fmlh::ExpPair kExp[128] = {FMLH_EXP2_128_PAIRS};
static const fmlh::SinCosEntry kSin[116] = {FMLH_SINCOS_128};

static double ulp_err(double v, long double ref) {
  const double rd = (double)ref;
  if (rd == 0.0) return v == 0.0 ? 0.0 : 1e9;
  const double u = std::nextafter(std::fabs(rd), INFINITY) - std::fabs(rd);
  return (double)(std::fabs((long double)v - ref) / (long double)u);
}

struct Stats {
  const char *name;
  long count = 0, libm_diff = 0;
  double worst = 0.0, wx = 0.0;
  void add(double x, double v, long double ref, double libm) {
    ++count;
    const double e = ulp_err(v, ref);
    if (e > worst) {
      worst = e;
      wx = x;
    }
    if (v != libm) ++libm_diff;
  }
  void print() const { std::printf("%s %ld %.6f %.17g %ld\n", name, count, worst, wx, libm_diff); }
};

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  std::mt19937_64 rng(12345);

  Stats se{"exp_neg"};
  auto check = [&](double x) {
    if (!(x > -16.0) || x > 0.0) return;
    se.add(x, fmlh::exp_neg(x, kExp), expl((long double)x), std::exp(x));
  };
  std::uniform_real_distribution<double> u16(-16.0, 0.0);
  for (long i = 0; i < n; ++i) check(u16(rng));
  // both sides of every argument-reduction boundary
  const double step = 0.6931471805599453 / 128.0;
  for (int k = 0; k >= -2960; --k) {
    for (double c : {(k + 0.5) * step, k * step}) {
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
    }
  }
  check(-0.0);
  check(0.0);
  check(std::nextafter(-16.0, 0.0));
  se.print();

  Stats sw{"exp_libm"};
  auto check_w = [&](double x) { sw.add(x, fmlh::exp_libm(x, kExp), expl((long double)x), std::exp(x)); };
  std::uniform_real_distribution<double> u760(-760.0, 0.0);
  for (long i = 0; i < n; ++i) check_w(u760(rng));
  for (double c : {-512.0, -707.0, -708.3964185322641, -745.1332191019411, -1024.0}) {
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
  for (double x = 1e-300; x < 1e-9; x *= 3.0) {
    check_w(x);
    check_w(-x);
  }
  sw.print();

  const double hp = 1.5707963267948966;
  Stats ss{"sin_libm"};
  auto check_s = [&](Stats &st, double y) {
    st.add(y, fmlh::sin_libm(y, kSin), sinl((long double)y), std::sin(y));
  };
  std::uniform_real_distribution<double> usin(0.126, hp);
  for (long i = 0; i < n; ++i) check_s(ss, usin(rng));
  for (int i = 16; i <= 205; ++i)
    for (double c : {(i + 0.5) / 128.0, i / 128.0, 0.85546875, hp}) {
      double y = c;
      for (int s = 0; s < 16; ++s) {
        if (y >= 0.126 && y <= hp + 1e-15) check_s(ss, y);
        y = std::nextafter(y, 2.0);
      }
      y = c;
      for (int s = 0; s < 16; ++s) {
        y = std::nextafter(y, 0.0);
        if (y >= 0.126 && y <= hp + 1e-15) check_s(ss, y);
      }
    }
  // the ARAD domain: y = d * (pi / (2 cut)) for d from 0.126 (2 cut)/pi to cut
  for (double cut : {5.0, 6.0, 3.0}) {
    const double ic = (4.0 * std::atan(1.0)) / (2.0 * cut);
    std::uniform_real_distribution<double> ud(0.126 / ic, cut);
    for (long i = 0; i < n / 4; ++i) check_s(ss, ud(rng) * ic);
    check_s(ss, cut * ic);
  }
  ss.print();

  Stats st{"sin_libm_small"};
  std::uniform_real_distribution<double> usmall(0.0, 0.126);
  for (long i = 0; i < n / 4; ++i) check_s(st, usmall(rng));
  for (double y = 1e-300; y < 0.126; y *= 1.7) check_s(st, y);
  check_s(st, 0.0);
  st.print();

  Stats sc{"sin_cr"};
  std::uniform_real_distribution<double> uhp(0.0, hp);
  for (long i = 0; i < n; ++i) {
    const double y = uhp(rng);
    sc.add(y, fmlh::sin_cr(y), sinl((long double)y), std::sin(y));
  }
  sc.print();
  return 0;
}
