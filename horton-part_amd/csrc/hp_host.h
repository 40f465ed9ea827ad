// Host-side helpers shared by the translation units of the library: the
// thread-local error text behind hp_last_error(), HIP error checks and the
// launch-shape rules of the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hortonpart.h"
#include "hp_device.h"

namespace hp {

inline thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

template <typename T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(HP_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

// Points per thread of the weights pass.  Every atom record is a broadcast
// LDS read shared by a thread's points: with one point per thread the reads
// of the four SIMDs nearly saturate the CU's LDS, so shells of more than one
// wavefront of points take two points per thread.
inline int weights_pt(int maxL) { return maxL > 64 ? 2 : 1; }

// Workgroup size of the weights pass for a molecule whose largest radial
// shell has maxL points.
inline int weights_block(int maxL) {
  const int pt = weights_pt(maxL);
  int bs = ((maxL + pt - 1) / pt + 63) / 64 * 64;
  return bs < 64 ? 64 : (bs > 512 ? 512 : bs);
}

// Workgroup shape of the weights pass for a molecule whose largest radial
// shell has maxL points: pt points per thread, bs threads, and up to g
// consecutive shells of one atom per workgroup (g maxL <= pt bs).  A shell
// rarely fills whole wavefronts (194 points take 97 of 128 lanes at two per
// thread); g is picked for the fullest lanes, the smallest g among equals.
#define HP_WGROUP_MAX 16
struct WeightsShape {
  int pt, bs, g;
};
inline WeightsShape weights_shape(int maxL) {
  const int pt = weights_pt(maxL);
  WeightsShape best{pt, weights_block(maxL), 1};
  double best_fill = -1.0;
  for (int g = 1; g <= HP_WGROUP_MAX; ++g) {
    const int n = g * maxL, threads = (n + pt - 1) / pt;
    if (threads > 512) break;
    const int bs = threads <= 64 ? 64 : (threads + 63) / 64 * 64;
    const double fill = (double)n / (double)(pt * bs);
    if (fill > best_fill + 1e-9) {
      best_fill = fill;
      best = {pt, bs, g};
    }
  }
  return best;
}

// First radial shell of every weights workgroup: groups of g consecutive
// shells of one atom (the last group of an atom may be shorter).
// atom_shell_off: natom+1 shell offsets per atom.
inline std::vector<int> weights_groups(const std::vector<int> &atom_shell_off, int g) {
  std::vector<int> first;
  for (size_t a = 0; a + 1 < atom_shell_off.size(); ++a)
    for (int s = atom_shell_off[a]; s < atom_shell_off[a + 1]; s += g) first.push_back(s);
  return first;
}

// Dynamic LDS of a weights launch: the staged atom records, or at the end two
// doubles per point of the workgroup, whichever is larger.
inline size_t weights_lds(size_t records, const WeightsShape &w) {
  const size_t pts = 2 * sizeof(double) * (size_t)w.pt * (size_t)w.bs;
  return (records > pts ? records : pts) + 16;
}

// MBIS solver layout of a molecule whose atoms have at most maxR radial
// points: 2 for one wavefront per atom with two points per lane (up to 128
// points; measured faster than two wavefronts and a barrier per iteration),
// else 1 for one workgroup per atom with one point per lane (three points per
// lane measured no faster at 150).  It depends on the molecule alone, so that
// its sums are the same alone and in any batch.
inline int mbis_ppl(int maxR) { return maxR <= 128 ? 2 : (maxR <= 192 ? 3 : (maxR <= 256 ? 4 : 1)); }

// Radial points per lane of the one-wave-per-atom solvers.
inline int pick_ppl(int R) {
  if (R <= 128) return 2;
  if (R <= 256) return 4;
  if (R <= 512) return 8;
  return 0;
}

// Sum the spread evaluation counters of the weights pass (HP_EVAL_SLOTS
// device values) into [FP32, FP64] totals.
inline int read_eval_counters(const unsigned long long *dev, double *fp32, double *fp64) {
  static_assert(HP_EVAL_SLOTS == 2 * HP_EVAL_LANES * HP_EVAL_STRIDE, "counter layout");
  unsigned long long h[HP_EVAL_SLOTS];
  hipError_t e = hipMemcpy(h, dev, sizeof(h), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(HP_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  unsigned long long t[2] = {0, 0};
  for (int p = 0; p < 2; ++p)
    for (int l = 0; l < HP_EVAL_LANES; ++l) t[p] += h[((size_t)p * HP_EVAL_LANES + l) * HP_EVAL_STRIDE];
  *fp32 = (double)t[0];
  *fp64 = (double)t[1];
  return 0;
}

}  // namespace hp

#define HP_HIP(expr)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return hp::fail(HP_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

#define HP_LAUNCHED()                                                                                     \
  do {                                                                                                    \
    hipError_t e_ = hipGetLastError();                                                                    \
    if (e_ != hipSuccess) return hp::fail(HP_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)
