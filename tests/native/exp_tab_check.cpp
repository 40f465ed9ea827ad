// Accuracy of the FP64 table exponential of the weights pass (hp_exp_tab.h)
// against long-double exp: prints the largest error in units in the last
// place over a sweep of [-745, 0] plus the special points, and the count of
// arguments off by more than one ulp.
#include <cmath>
#include <cstdio>
using std::fma;
using std::fmax;
using std::ldexp;
using std::rint;
#include "hp_exp_tab.h"

int main() {
  static const double t64[64] = HP_EXP2_64_VALUES;
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
