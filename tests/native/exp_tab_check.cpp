// Accuracy of the FP64 table exponentials (hp_exp_tab.h) against long double:
// exp(x) over a sweep of [-745, 0] plus special points, and 2^(y/64) (the
// scaled form of the weights pass) over the same range in units of ln2/64
// against exp2l(y/64) (y/64 is exact).  Prints per function the largest error
// in units in the last place and the count of arguments off by more than one
// ulp.
#include <cmath>
#include <cstdio>
using std::fma;
using std::fmax;
using std::ldexp;
using std::rint;
#include "hp_exp_tab.h"

static const double t64[64] = HP_EXP2_64_VALUES;

// 2^(y/64) against exp2l: same error measure as below
static void check_scaled() {
  double worst = 0.0, worst_y = 0.0;
  long over1 = 0, n = 0;
  auto check = [&](double y) {
    const double got = hp_exp2_64th_core(y, t64);
    const long double ref = exp2l((long double)y / 64.0L);
    if (ref < 2.2250738585072014e-308L) {
      const long double err = fabsl((long double)got - ref) / 4.9406564584124654e-324L;
      if ((double)err > worst && ref > 0) { worst = (double)err; worst_y = y; }
      ++n;
      return;
    }
    const double ulp = std::nextafter((double)ref, INFINITY) - (double)ref;
    const double e = (double)(fabsl((long double)got - ref) / (long double)ulp);
    if (e > worst) { worst = e; worst_y = y; }
    if (e > 1.0) ++over1;
    ++n;
  };
  const double ymin = -745.0 * HP_64_OVER_LN2;
  for (long i = 0; i <= 4000000; ++i) check(ymin * (double)i / 4000000.0);
  for (long i = 0; i <= 1000000; ++i) check(-64.0 * (double)i / 1000000.0);
  const double specials[] = {0.0, -0.0, -1e-300, -0.5, -31.5, -32.5, -63.5, -64.0, -1e5, -1e9};
  for (double y : specials) check(y);
  printf("scaled: n=%ld max_ulp=%.3f at y=%.17g over_1ulp=%ld\n", n, worst, worst_y, over1);
  printf("scaled: 2^0=%.17g 2^(-1e9/64)=%.17g\n", hp_exp2_64th_core(0.0, t64), hp_exp2_64th_core(-1e9, t64));
}

int main() {
  check_scaled();
  double worst = 0.0, worst_x = 0.0;
  long over1 = 0, n = 0;
  auto check = [&](double x) {
    const double got = hp_exp_tab_core(x, t64);
    const long double ref = expl((long double)x);
    if (ref < 2.2250738585072014e-308L) {  // subnormal or zero result: absolute check
      const long double err = fabsl((long double)got - ref) / 4.9406564584124654e-324L;
      if ((double)err > worst && ref > 0) { worst = (double)err; worst_x = x; }
      ++n;
      return;
    }
    const double ulp = std::nextafter((double)ref, INFINITY) - (double)ref;
    const double e = (double)(fabsl((long double)got - ref) / (long double)ulp);
    if (e > worst) { worst = e; worst_x = x; }
    if (e > 1.0) ++over1;
    ++n;
  };
  for (long i = 0; i <= 4000000; ++i) check(-745.0 * (double)i / 4000000.0);
  for (long i = 0; i <= 1000000; ++i) check(-1.0 * (double)i / 1000000.0);
  const double specials[] = {0.0, -0.0, -1e-300, -1e-17, -0.0054152123481245725, -0.69314718055994529, -708.39641853226408,
                             -744.0, -745.1332191019411, -746.0, -1000.0, -1e6};
  for (double x : specials) check(x);
  printf("n=%ld max_ulp=%.3f at x=%.17g over_1ulp=%ld\n", n, worst, worst_x, over1);
  printf("exp(0)=%.17g exp(-1e6)=%.17g\n", hp_exp_tab_core(0.0, t64), hp_exp_tab_core(-1e6, t64));
  return 0;
}
