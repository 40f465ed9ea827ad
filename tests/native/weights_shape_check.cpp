// Invariants of the weights pass's workgroup shapes and shell groups
// (hp_host.h), which the kernel relies on: every point of a group has a
// thread slot (g maxL <= pt bs), the shape fills the lanes at least as well
// as one shell per workgroup, and the groups cover every radial shell of
// every atom once, in order, without crossing an atom.  Prints "ok" or the
// first failure.
#include <cstdio>
#include <random>
#include <vector>

#include "hp_host.h"

static int fail(const char *what, int a, int b) {
  printf("FAIL %s (%d, %d)\n", what, a, b);
  return 1;
}

int main() {
  for (int maxL = 1; maxL <= 1024; ++maxL) {
    const hp::WeightsShape w = hp::weights_shape(maxL);
    if (!(w.pt == 1 || w.pt == 2)) return fail("points per thread", maxL, w.pt);
    if (w.bs < 64 || w.bs > 512 || w.bs % 64) return fail("block size", maxL, w.bs);
    if (w.g < 1 || w.g > HP_WGROUP_MAX) return fail("shells per group", maxL, w.g);
    if ((long)w.g * maxL > (long)w.pt * w.bs) return fail("points beyond the thread slots", maxL, w.g);
    const int bs1 = hp::weights_block(maxL);
    const double fill1 = (double)maxL / (w.pt * bs1), fill = (double)w.g * maxL / (w.pt * w.bs);
    if (fill + 1e-12 < fill1) return fail("fill below one shell per workgroup", maxL, w.g);
  }
  const int want[][4] = {{194, 2, 512, 5}, {302, 2, 320, 2}, {110, 2, 448, 8}, {26, 1, 320, 12}};
  for (const auto &c : want) {
    const hp::WeightsShape w = hp::weights_shape(c[0]);
    if (w.pt != c[1] || w.bs != c[2] || w.g != c[3]) return fail("expected shape", c[0], w.g);
  }
  std::mt19937 rng(7);
  for (int trial = 0; trial < 2000; ++trial) {
    const int natom = 1 + (int)(rng() % 40), g = 1 + (int)(rng() % HP_WGROUP_MAX);
    std::vector<int> aso(natom + 1, 0);
    for (int a = 0; a < natom; ++a) aso[a + 1] = aso[a] + 1 + (int)(rng() % 200);
    const std::vector<int> first = hp::weights_groups(aso, g);
    // the kernel's group length: min(g, end of the first shell's atom - first)
    std::vector<int> covered(aso[natom], 0);
    int atom = 0, prev = -1;
    for (int s : first) {
      if (s <= prev) return fail("group starts not increasing", trial, s);
      prev = s;
      while (aso[atom + 1] <= s) ++atom;
      const int gn = std::min(g, aso[atom + 1] - s);
      for (int k = s; k < s + gn; ++k) covered[k] += 1;
    }
    for (int k = 0; k < aso[natom]; ++k)
      if (covered[k] != 1) return fail("shell not covered exactly once", trial, k);
  }
  printf("ok\n");
  return 0;
}
