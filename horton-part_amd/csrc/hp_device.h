// Device code of the stockholder-partitioning library (gfx950, wave64).
//
// Kernels
//   k_iter_begin      pro-atom function table (coefficient, exponent) from the
//                     current parameters (mbis.py:284-289, basis.py:161-171)
//   k_weights<R,F,G>  fused pass over a group of radial shells per workgroup:
//                     pro-atoms of every (or every neighbouring) atom,
//                     promolecule with the +1e-100 floor, own-atom weight
//                     clipped to [0,1], spherical average of w*rho and the
//                     entropy partial (stockholder.py:145-175, 352-384;
//                     mbis.py:176-184).  R = float gives the mixed path.
//   k_mbis_inner      one workgroup per atom: MBIS fixed point (mbis.py:81-163)
//   k_gisa_qp         one wavefront per atom: QP over the simplex (gisa.py:348-421)
//   k_lisa_inner      one wavefront per atom: sc / newton / m-newton on the
//                     radial grid (alisa.py:193-291, 753-938)
//   k_finalize        entropy and change sums, histories (iterstock.py:32-58)
//   k_moments*        AIM multipoles (core/base.py:329-402)
//   k_batch_*         the outer-loop bookkeeping of many molecules at once
//                     (one slot record per molecule in flight)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HP_HD __host__ __device__
#include "hp_exp_tab.h"
#include "hp_spline.h"
#include "hp_moments.h"

#define HP_KMAX 16      // functions per atom handled by the per-atom kernels

// per-atom status bits (same values as HP_ST_* in include/hortonpart.h)
#define HP_ST_NOCONV 1
#define HP_ST_POP 2
#define HP_ST_NEGDENS 4
#define HP_ST_LINESEARCH 8
#define HP_ST_NEWTONMAX 16
#define HP_ST_QPFAIL 32
#define HP_ST_NEGPARS 64
#define HP_ST_NOTMONO 128
#define HP_ST_SINGULAR 256
#define HP_ST_NONFINITE 512
#define HP_ST_NORECORD 1024
#define HP_PI 3.14159265358979323846
// The FP32 stage of the mixed path ends when the change is below
// threshold * HP_MIXED_MARGIN: just above the final test, so that the FP64
// redo of that iteration starts no later than the plain FP64 path stops.
#define HP_MIXED_MARGIN 1.01

// status bits that end the outer loop (the reference raises on them)
#define HP_ST_FATAL                                                                                         \
  (HP_ST_NEGDENS | HP_ST_LINESEARCH | HP_ST_NEWTONMAX | HP_ST_QPFAIL | HP_ST_NOTMONO | HP_ST_SINGULAR | \
   HP_ST_NONFINITE | HP_ST_NORECORD)

// Stage of a molecule in flight (HpSlot::run)
#define HP_RUN_IDLE 0  // finished, failed, or waiting for the host
#define HP_RUN_FP32 1  // mixed precision: FP32 grid arithmetic
#define HP_RUN_FP64 2  // plain FP64 (or the FP64 polish of the mixed path)

// Why a molecule stopped (HpSlot::event)
#define HP_EV_RUNNING 0
#define HP_EV_CONVERGED 1  // change < threshold
#define HP_EV_MAXITER 2    // iteration cap reached first
#define HP_EV_ERROR 3      // a per-atom solver raised
#define HP_EV_SWITCH 4     // FP32 stage done: the host moves it to FP64

// Outer-loop state of one molecule in a batch, owned by the device between
// host polls.  The host fills it at admission and flips run from IDLE to
// FP64 after an HP_EV_SWITCH.
struct HpSlot {
  int run, event;
  int scheme;          // 0: MBIS (N, S pairs), 1: linear coefficients
  int a0, na;          // atoms [a0, a0 + na) of the batch arrays
  int s0, ns;          // radial shells
  int f0, nf;          // pro-atom functions
  int niter;           // outer iterations done
  int min_stop, max_iter;  // bounds of the current stage
  int n_polish, maxiter;   // the request's settings
  int err, any;        // status bits: fatal ones of the last iteration, OR of all
  int corr_period;     // FP32 stage: FP64 refresh of the correction every this many iterations (0: none)
  int refresh;         // this iteration refreshes the correction (set by k_batch_begin)
  int full_sum;        // weights pass: every atom on every point (no tail skipping)
  int phase;           // 0: next tick starts an outer iteration; 1: its inner solves continue
  int wgroup;          // weights pass: radial shells per workgroup
  int pad_;
  double thr;          // threshold of the current stage
  double threshold;    // the request's threshold
  double change, entropy;
};

struct DevMol {
  int natom;
  int nshell;
  int64_t npoint;
  const double *cx, *cy, *cz;        // atom centres
  const double *px, *py, *pz;        // grid points (SoA)
  const double *rho, *wmol, *wat, *wsph;
  const int *shell_atom;
  const int64_t *shell_off;
  const double *shell_r, *shell_w;
  const int *atom_shell_off;         // natom+1 radial-shell offsets per atom
  const double *pseudo;
  // Batched form (many molecules in one launch; both null for one molecule):
  // work[i] = (shell or atom index, slot of its molecule) for block/wave i,
  // and natom then counts the entries of an atom work list.
  const int2 *work;
  const HpSlot *slots;
  // Weights pass: a workgroup covers up to wgroup consecutive radial shells
  // of one atom (batched: HpSlot::wgroup); grp_first[i] is the first shell of
  // workgroup i of one molecule (null: one shell per workgroup, in order).
  const int *grp_first;
  int wgroup;
};

struct DevFn {
  const int *fn_off;     // natom+1
  const double *coef;    // per function: prefactor (population included)
  const double *alpha;   // per function: exponent
  const double *order;   // per function: power of r
  const int *nb_list;    // natom*natom neighbour lists sorted by distance (culling)
  const int *shell_nb;   // nshell: candidate count of each shell (culling)
  const int64_t *nb_off; // per atom: start of its list (null: atom * natom)
  double rc;             // cutoff radius; <= 0 disables culling
  int tail_cull;         // weights pass: skip atoms whose bounded contribution is
                         // negligible (schemes with non-negative coefficients)
  unsigned long long *evals_done;  // HP_EVAL_SLOTS counters per precision (may be null)
  // tabulated pro-atoms (HP_FN_TABLE): atom b's knots tab_x[tab_off[b] ..
  // tab_off[b+1]) and 4 cubic coefficients per knot in tab_c
  const int *tab_off = nullptr;
  const double *tab_x = nullptr;
  const double *tab_c = nullptr;
  // interval guide: per atom HP_TAB_GUIDE start intervals over uniform bins
  // of log2 r, and (log2 of the first knot, bins per unit of log2 r)
  const int *tab_guide = nullptr;
  const float2 *tab_gparam = nullptr;
};

#define HP_TAB_GUIDE 256

// Counters of performed function evaluations: 64 per precision, one per
// 128-byte line, picked by workgroup, so that the workgroups of a launch do
// not all add to one address (the host sums them).
#define HP_EVAL_LANES 64
#define HP_EVAL_STRIDE 16  // unsigned long longs per 128-byte line
#define HP_EVAL_SLOTS (2 * HP_EVAL_LANES * HP_EVAL_STRIDE)

// Function kinds of the promolecule pass.
#define HP_FN_SLATER 0   // all functions exp(-a r)
#define HP_FN_GAUSS 1    // all functions exp(-a r^2)
#define HP_FN_GENERAL 2  // exp(-a r^n), n per function
#define HP_FN_TABLE 3    // piecewise cubic radial table per atom (ISA, Hirshfeld, Hirshfeld-I)

// ---------------------------------------------------------------------------
// reductions (deterministic: fixed tree, result broadcast from lane 0)
// ---------------------------------------------------------------------------

// One DPP step: the value of lane (pattern CTRL) for lanes of the rows in
// ROWMASK, 0 elsewhere.  Doubles move as two 32-bit halves.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane63(double v) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Wave-wide sums with DPP row operations: pairs, quads, half rows, rows, then
// rows 0+1 and 0..3 by row broadcasts; lane 63 ends with the total, which is
// read into a scalar register so that every lane sees the same value.  The
// tree is fixed, so results are deterministic.  All 64 lanes must be active.
template <int N>
__device__ __forceinline__ void wave_sum_n(double *v) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_d<0xB1, 0xF>(v[i]);  // quad_perm [1,0,3,2]
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_d<0x4E, 0xF>(v[i]);  // quad_perm [2,3,0,1]
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_d<0x141, 0xF>(v[i]);  // row_half_mirror
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_d<0x140, 0xF>(v[i]);  // row_mirror
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_d<0x142, 0xA>(v[i]);  // row_bcast:15 -> rows 1, 3
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_d<0x143, 0xC>(v[i]);  // row_bcast:31 -> rows 2, 3
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = lane63(v[i]);
}

// All-reduce of N doubles over a wavefront, for latency-bound loops: DPP
// butterflies inside each row of 16 lanes (pairs, quads, half rows, rows),
// then the gfx950 lane swaps across rows, 32 lanes apart and then 16 apart.
// Called with one register as both operands, a swap returns the lane's own
// value and its partner's, whose sum is the pair total; each add sees its
// two operands in the same order in every lane, so all 64 lanes end with the
// same total, bit for bit.  All 64 lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_row_d(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double swap32_sum(double v) {
  const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
  const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

__device__ __forceinline__ double swap16_sum(double v) {
  const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
  const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

template <int N>
__device__ __forceinline__ void wave_allsum_n(double *v) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_row_d<0xB1>(v[i]);  // quad_perm [1,0,3,2]
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_row_d<0x4E>(v[i]);  // quad_perm [2,3,0,1]
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_row_d<0x141>(v[i]);  // row_half_mirror
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += dpp_row_d<0x140>(v[i]);  // row_mirror
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = swap32_sum(v[i]);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = swap16_sum(v[i]);
}

// Four sums in one register.  v_permlane32_swap exchanges lanes 32-63 of its
// first operand with lanes 0-31 of its second, so for two values a and b the
// sum of the swapped pair holds a_i + a_(i+32) in lanes 0-31 and
// b_i + b_(i+32) in lanes 32-63.  v_permlane16_swap exchanges the odd rows
// (16 lanes) of its first operand with the even rows of its second: on two
// such registers (a|b and c|d) the swapped sum holds the quarter sums of a,
// c, b, d in rows 0, 1, 2, 3, and the row reductions finish all four at once.
__device__ __forceinline__ double pair_swap32_sum(double a, double b) {
  const auto l = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false,
                                                  false);
  const auto h = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false,
                                                  false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

__device__ __forceinline__ double pair_swap16_sum(double p, double q) {
  const auto l = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(p), (unsigned)__double2loint(q), false,
                                                  false);
  const auto h = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(p), (unsigned)__double2hiint(q), false,
                                                  false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

__device__ __forceinline__ double lane_value(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Wave-wide sums of N values (every lane of the wavefront active), four per
// register through the swaps above, then the row reductions; the sums come
// back as wave-uniform values.  Fewer instructions than N separate trees:
// the swaps need no copies, one addition serves two values.
template <int N>
__device__ __forceinline__ void wave_allsum_packed(double *v) {
  constexpr int G = (N + 3) / 4;
  double r[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int i = 4 * g;
    const double p = i + 1 < N ? pair_swap32_sum(v[i], v[i + 1]) : pair_swap32_sum(v[i], 0.0);
    if (i + 2 < N) {
      const double q = i + 3 < N ? pair_swap32_sum(v[i + 2], v[i + 3]) : pair_swap32_sum(v[i + 2], 0.0);
      r[g] = pair_swap16_sum(p, q);
    } else {
      r[g] = pair_swap16_sum(p, 0.0);
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) r[g] += dpp_row_d<0xB1>(r[g]);  // quad_perm [1,0,3,2]
#pragma unroll
  for (int g = 0; g < G; ++g) r[g] += dpp_row_d<0x4E>(r[g]);  // quad_perm [2,3,0,1]
#pragma unroll
  for (int g = 0; g < G; ++g) r[g] += dpp_row_d<0x141>(r[g]);  // row_half_mirror
#pragma unroll
  for (int g = 0; g < G; ++g) r[g] += dpp_row_d<0x140>(r[g]);  // row_mirror
  // rows 0, 1, 2, 3 hold values 4g, 4g + 2, 4g + 1, 4g + 3
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int i = 4 * g;
    v[i] = lane_value(r[g], 0);
    if (i + 1 < N) v[i + 1] = lane_value(r[g], 32);
    if (i + 2 < N) v[i + 2] = lane_value(r[g], 16);
    if (i + 3 < N) v[i + 3] = lane_value(r[g], 48);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  wave_sum_n<1>(&v);
  return v;
}

// Wave-wide minimum (same DPP tree; masked rows contribute +inf).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_min_step(double v) {
  const double inf = __builtin_inf();
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(inf), __double2loint(v), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(inf), __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  return fmin(v, __hiloint2double(hi, lo));
}

__device__ __forceinline__ double wave_min(double v) {
  v = dpp_min_step<0xB1, 0xF>(v);
  v = dpp_min_step<0x4E, 0xF>(v);
  v = dpp_min_step<0x141, 0xF>(v);
  v = dpp_min_step<0x140, 0xF>(v);
  v = dpp_min_step<0x142, 0xA>(v);
  v = dpp_min_step<0x143, 0xC>(v);
  return lane63(v);
}

// Block sum of two values (blockDim multiple of 64, <= 1024); thread 0 gets it.
__device__ __forceinline__ void block_sum2(double &a, double &b, double *lds) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if (lane == 0) {
    lds[2 * wv] = a;
    lds[2 * wv + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int w = 0; w < nw; ++w) {
      sa += lds[2 * w];
      sb += lds[2 * w + 1];
    }
    a = sa;
    b = sb;
  }
}

template <typename R>
__device__ __forceinline__ R hp_exp(R x);
template <>
__device__ __forceinline__ double hp_exp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float hp_exp<float>(float x) { return __expf(x); }

template <typename R>
__device__ __forceinline__ R hp_sqrt(R x);
template <>
__device__ __forceinline__ double hp_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float hp_sqrt<float>(float x) { return __fsqrt_rn(x); }

// np.clip(w, 0, 1): NaN stays NaN, +-inf go to the bounds.
__device__ __forceinline__ double clip01(double w) { return w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w); }

// AIM weight own/promol (stockholder.py:370-377).  In FP64 the promolecule
// carries the 1e-100 floor, so the division is the reference's; the FP32
// promolecule has no floor and an exact zero gives weight 0.
__device__ __forceinline__ double aim_weight(double own, double promol) { return clip01(own / promol); }
__device__ __forceinline__ double aim_weight(float own, float promol) {
  return promol != 0.0f ? clip01((double)(own / promol)) : 0.0;
}

template <typename R>
__device__ __forceinline__ R hp_floor_add();
template <>
__device__ __forceinline__ double hp_floor_add<double>() { return 1e-100; }
template <>
__device__ __forceinline__ float hp_floor_add<float>() { return 0.0f; }

// ---------------------------------------------------------------------------
// function table
// ---------------------------------------------------------------------------

// Start of an outer iteration: keep the parameters that produce this
// iteration's weights (old_par) and build the pro-atom function table.
// MBIS: par[2f] = N, par[2f+1] = S -> coef = N S^3 / (8 pi), alpha = S.
// Linear bases: coef = c_f * norm_f, alpha fixed.
static __global__ void k_iter_begin(const int *__restrict__ done, int scheme, int npar, int nfn, const double *__restrict__ par,
                             double *__restrict__ old_par, const double *__restrict__ norm,
                             const double *__restrict__ alpha_fixed, double *__restrict__ coef,
                             double *__restrict__ alpha) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npar) old_par[i] = par[i];
  if (i < nfn) {
    if (scheme == 0) {
      const double N = par[2 * i], S = par[2 * i + 1];
      coef[i] = N * (S * S * S) / (8.0 * HP_PI);
      alpha[i] = S;
    } else if (scheme == 2) {
      // NLIS/GMBIS (N, S, n): N n S^(3/n) / (4 pi Gamma(3/n)), norm = n / (4 pi Gamma(3/n))
      const double N = par[3 * i], S = par[3 * i + 1], n = par[3 * i + 2];
      coef[i] = N * norm[i] * pow(S, 3.0 / n);
      alpha[i] = S;
    } else {
      coef[i] = par[i] * norm[i];
      alpha[i] = alpha_fixed[i];
    }
  }
}

// (dx^2 + dy^2) + dz^2 with every operation rounded on its own.
__device__ __forceinline__ double dist2_rounded(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return (dx * dx + dy * dy) + dz * dz;
}

// Pro-atom density of atom b at squared distance r2 (and distance r).
template <typename R, int FK>
__device__ __forceinline__ R proatom_value(const DevFn &fn, int b, R r, R r2) {
  if (FK == HP_FN_TABLE) {
    // the spline plus the 1e-100 of eval_proatom (core/stockholder.py:347-349,
    // hirshfeld_i.py:158); the interval starts from the guide bin of log2 r
    // and walks to the exact one (the largest i <= n - 2 with x_i <= r, or 0)
    const int k0 = fn.tab_off[b], n = fn.tab_off[b + 1] - k0;
    const double *x = fn.tab_x + k0;
    const double rd = (double)r;
    const float2 gp = fn.tab_gparam[b];
    float t = (__log2f((float)rd) - gp.x) * gp.y;
    t = t > 0.0f ? t : 0.0f;  // below the first knot, r = 0 and NaN
    t = t < (float)(HP_TAB_GUIDE - 1) ? t : (float)(HP_TAB_GUIDE - 1);
    int i = fn.tab_guide[b * HP_TAB_GUIDE + (int)t];
    while (i < n - 2 && x[i + 1] <= rd) ++i;
    while (i > 0 && x[i] > rd) --i;
    return (R)(table_piece(fn.tab_c + 4 * (int64_t)(k0 + i), rd - x[i]) + 1e-100);
  }
  R pb = (R)0;
  const int f1 = fn.fn_off[b + 1];
  for (int f = fn.fn_off[b]; f < f1; ++f) {
    const R c = (R)fn.coef[f];
    const R al = (R)fn.alpha[f];
    R arg;
    if (FK == HP_FN_SLATER) {
      arg = r;
    } else if (FK == HP_FN_GAUSS) {
      arg = r2;
    } else {
      const double n = fn.order[f];
      arg = (n == 1.0) ? r : ((n == 2.0) ? r2 : (R)pow((double)r, n));
    }
    pb += c * hp_exp<R>(-al * arg);
  }
  return pb;
}

// Promolecule and own-atom pro-atom at one point (reference accumulation
// order: atom index order, +1e-100 after each atom, stockholder.py:167-170).
template <typename R, int FK>
__device__ __forceinline__ void promol_at(const DevMol &m, const DevFn &fn, int a, int nb_n, const int *nb, double x,
                                          double y, double z, R &promol, R &own) {
  promol = (R)0;
  own = (R)0;
  const bool cull = fn.rc > 0.0;
  const R rc = (R)fn.rc;
  for (int t = 0; t < nb_n; ++t) {
    const int b = nb ? nb[t] : t;
    const double dx = x - m.cx[b], dy = y - m.cy[b], dz = z - m.cz[b];
    // tabulated pro-atoms: the distance as numpy's norm forms it
    const R r2 = (R)(FK == HP_FN_TABLE ? dist2_rounded(dx, dy, dz) : dx * dx + dy * dy + dz * dz);
    const R r = (FK == HP_FN_GAUSS && !cull) ? (R)0 : hp_sqrt<R>(r2);
    R pb = (R)0;
    if (!cull || r <= rc) pb = proatom_value<R, FK>(fn, b, r, r2);
    promol += pb;
    promol += hp_floor_add<R>();
    if (b == a) own = pb;
  }
  if (nb) {
    // culled atoms still contribute their 1e-100 floor
    for (int t = nb_n; t < m.natom; ++t) promol += hp_floor_add<R>();
  }
}

// ---------------------------------------------------------------------------
// K1: fused weights pass, one workgroup per radial shell.
// The atoms that can reach the shell (all atoms of its molecule, or the
// neighbour list when a cut-off is set) are visited in tiles staged in LDS as
// fixed-size records {x, y, z, nf, coef[KM], alpha[KM]}, read by every lane
// as a broadcast (one wide read per record).  Each thread owns PT points.
// FP64 path: distances from absolute coordinates, exp().  FP32 path:
// coordinates relative to the shell's atom (offsets formed in FP64), raw
// v_sqrt_f32, and exponents pre-scaled by log2(e) so that each function is one
// multiply and one v_exp_f32.  The promolecule is accumulated in atom order
// with the reference's 1e-100 floor after every atom (FP64 only: in FP32 the
// floor rounds away).
// ---------------------------------------------------------------------------
template <typename R>
__host__ __device__ constexpr int weights_record(int km) {
  return 4 + 2 * km;
}
template <typename R>
__host__ __device__ constexpr size_t weights_lds_per_atom(int km) {
  return (size_t)weights_record<R>(km) * sizeof(R) + sizeof(int);
}

// Record bytes of a weights launch's staging rounds: `tile` records in R
// and, for the FP64 pass of Slater or Gaussian pro-atoms (which can take the
// tail path), `tile` FP32 records of the FP32 tier as well.
template <typename R>
__host__ __device__ constexpr size_t weights_lds_records(int km, int tile, int fk) {
  return weights_lds_per_atom<R>(km) * (size_t)tile +
         (sizeof(R) == sizeof(double) && fk != HP_FN_GENERAL
              ? (size_t)weights_record<float>(km) * sizeof(float) * (size_t)tile
              : 0);
}

// Functions per atom record: the smallest of 2, 4, 8, 16 that holds the
// largest atom (general orders always use 16).
inline __host__ int weights_km(int maxK, int fk) {
  if (fk == HP_FN_GENERAL) return 16;
  return maxK <= 2 ? 2 : (maxK <= 4 ? 4 : (maxK <= 8 ? 8 : 16));
}

// Atoms staged in LDS at once (records of km functions in R, <= 40 KB).
template <typename R>
inline __host__ int weights_tile_size(int km, int natom_max) {
  const size_t per_atom = weights_lds_per_atom<R>(km);
  const int tile = (int)((40 * 1024) / per_atom < 256 ? (40 * 1024) / per_atom : 256);
  return tile < 1 ? 1 : (tile < natom_max ? tile : (natom_max < 1 ? 1 : natom_max));
}

// FP64 exp of the weights pass: Tang's table method (hp_exp_tab.h), with the
// 64 table values staged in LDS by each workgroup.
static __constant__ double hp_exp2_64[64] = HP_EXP2_64_VALUES;

__device__ __forceinline__ double exp_tab(double x, const double *__restrict__ t64) { return hp_exp_tab_core(x, t64); }

template <typename R>
__device__ __forceinline__ R fast_sqrt(R x);
// FP64: v_rsq_f64 refined by the Newton-Raphson steps of the correctly
// rounded sqrt, without its rescaling of inputs below 2^-767 (a squared
// distance is either 0 or far above that); 0 gives 0.
template <>
__device__ __forceinline__ double fast_sqrt<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y, h = 0.5 * y;
  const double e = fma(-h, s, 0.5);
  s = fma(s, e, s);
  h = fma(h, e, h);
  double d = fma(-s, s, x);
  s = fma(d, h, s);
  d = fma(-s, s, x);
  s = fma(d, h, s);
  return x > 0.0 ? s : 0.0;
}
template <>
__device__ __forceinline__ float fast_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// Distances of the weights pass.  FP64: v_rsq_f64, one Goldschmidt step and
// one Newton-Raphson step (the correctly rounded sequence above stops one
// step earlier): within about an ulp.  An ulp of r moves exp(-alpha r) by
// alpha r 2^-53 relative, as the rounding of r itself does.
template <typename R>
__device__ __forceinline__ R dist_sqrt(R x);
template <>
__device__ __forceinline__ double dist_sqrt<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y, h = 0.5 * y;
  const double e = fma(-h, s, 0.5);
  s = fma(s, e, s);
  h = fma(h, e, h);
  const double d = fma(-s, s, x);
  s = fma(d, h, s);
  return x > 0.0 ? s : 0.0;
}
template <>
__device__ __forceinline__ float dist_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// exp of -(alpha * arg) from the exponent as staged: FP64 by the table method
// with alpha pre-scaled by 64/ln2; FP32 exp2 with alpha pre-scaled by log2(e)
template <typename R>
__device__ __forceinline__ R neg_exp_scaled(R al, R arg, const double *t64);
template <>
__device__ __forceinline__ double neg_exp_scaled<double>(double al, double arg, const double *t64) {
  return hp_exp2_64th_core(-al * arg, t64);
}
template <>
__device__ __forceinline__ float neg_exp_scaled<float>(float al, float arg, const double *) {
  return __builtin_amdgcn_exp2f(-al * arg);
}

// Argument of function k of a record: r (Slater), r^2 (Gauss), r^n (general).
template <typename R, int FK>
__device__ __forceinline__ R fn_arg(R r, R r2, const DevFn &fn, int b, int k) {
  if (FK == HP_FN_SLATER) return r;
  if (FK == HP_FN_GAUSS) return r2;
  const double n = fn.order[fn.fn_off[b] + k];
  return (n == 1.0) ? r : ((n == 2.0) ? r2 : (R)pow((double)r, n));
}

// The atoms of one staged tile on the thread's first NP point slots.  The
// distances come first; then the functions, two at a time, each uniform
// branch holding 2 x NP independent exponentials.  Per point the terms are
// added in function order.  With the cut-off, an atom beyond it from every
// point of the wavefront is not evaluated (it adds its floors only, as the
// evaluated and zeroed terms would).  Returns the number of functions
// evaluated (uniform in the wavefront).
template <typename R, int FK, int PT, int NP, int KM, bool CULL>
__device__ __forceinline__ int weights_tile(const R *__restrict__ rec, const int *__restrict__ s_idx,
                                             const double *__restrict__ t64, int tn, int own_i,
                                             const DevFn &fn, R rc, const R (&qx)[PT], const R (&qy)[PT],
                                             const R (&qz)[PT], const double (&px)[PT], const double (&py)[PT],
                                             const double (&pz)[PT], R (&promol)[PT], R (&own)[PT]) {
  constexpr int RS = weights_record<R>(KM);
  int nfn = 0;
  for (int i = 0; i < tn; ++i) {
    const R *q = rec + (size_t)i * RS;
    R v[RS];
#pragma unroll
    for (int u = 0; u < RS; ++u) v[u] = q[u];
    const int nf = __builtin_amdgcn_readfirstlane((int)v[3]);
    const bool is_own = i == own_i;
    const int b = FK == HP_FN_GENERAL ? s_idx[i] : 0;
    R r[NP], r2[NP], pb[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      R dx, dy, dz;
      if (sizeof(R) == sizeof(double)) {
        dx = (R)(px[j] - (double)v[0]);
        dy = (R)(py[j] - (double)v[1]);
        dz = (R)(pz[j] - (double)v[2]);
      } else {
        dx = qx[j] - v[0];
        dy = qy[j] - v[1];
        dz = qz[j] - v[2];
      }
      r2[j] = dx * dx + dy * dy + dz * dz;
      r[j] = (FK == HP_FN_GAUSS && !CULL) ? (R)0 : dist_sqrt<R>(r2[j]);
      pb[j] = (R)0;
    }
    bool eval = true;
    if (CULL) {
      bool in = false;
#pragma unroll
      for (int j = 0; j < NP; ++j) in = in || r[j] <= rc;
      eval = __any(in);
    }
    if (eval) {
      nfn += nf;
#pragma unroll
      for (int k = 0; k < KM; k += 2) {
        if (k + 1 < nf) {
#pragma unroll
          for (int j = 0; j < NP; ++j) {
            pb[j] += v[4 + k] * neg_exp_scaled<R>(v[4 + KM + k], fn_arg<R, FK>(r[j], r2[j], fn, b, k), t64);
            pb[j] += v[5 + k] * neg_exp_scaled<R>(v[5 + KM + k], fn_arg<R, FK>(r[j], r2[j], fn, b, k + 1), t64);
          }
        } else if (k < nf) {
#pragma unroll
          for (int j = 0; j < NP; ++j)
            pb[j] += v[4 + k] * neg_exp_scaled<R>(v[4 + KM + k], fn_arg<R, FK>(r[j], r2[j], fn, b, k), t64);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      if (CULL && r[j] > rc) pb[j] = (R)0;
      promol[j] += pb[j];
      if (sizeof(R) == sizeof(double)) promol[j] += hp_floor_add<R>();
      if (is_own) own[j] = pb[j];
    }
  }
  return nfn;
}

// One atom record of the weights pass: position (FP64: absolute; FP32:
// relative to the shell's atom, formed in FP64), function count, prefactors
// and exponents (FP64: pre-scaled by 64/ln2; FP32: by log2 e).
template <typename R, int KM>
__device__ __forceinline__ void stage_record(R *__restrict__ q, int *__restrict__ idx, int b, const DevMol &m,
                                             const DevFn &fn, double ax, double ay, double az) {
  constexpr double log2e = 1.4426950408889634;
  *idx = b;
  if (sizeof(R) == sizeof(double)) {
    q[0] = (R)m.cx[b];
    q[1] = (R)m.cy[b];
    q[2] = (R)m.cz[b];
  } else {
    q[0] = (R)(m.cx[b] - ax);
    q[1] = (R)(m.cy[b] - ay);
    q[2] = (R)(m.cz[b] - az);
  }
  const int f0 = fn.fn_off[b], nf = fn.fn_off[b + 1] - f0;
  q[3] = (R)nf;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const bool on = k < nf;
    q[4 + k] = on ? (R)fn.coef[f0 + k] : (R)0;
    q[4 + KM + k] = on ? (R)(fn.alpha[f0 + k] * (sizeof(R) == sizeof(double) ? HP_64_OVER_LN2 : log2e)) : (R)0;
  }
}

// Bounded tails.  With non-negative coefficients (MBIS populations, GISA
// simplex coefficients) every pro-atom is a decreasing function of the
// distance.  For the points of one workgroup (radial shells of atom a), at
// distances up to rhi from R_a, atom b then contributes at most
//   U_b = sum_k c_bk f_k(x_b),   x_b = max(0, |R_b - R_a| - rhi),
// and the promolecule is at least L = sum_k c_ak f_k(rhi) + (floors).  An
// atom with U_b < eps L moves the promolecule of every point of the group by
// less than eps relative, so the pass may skip it (its 1e-100 floor is still
// added).  The test runs in the log2 domain in FP32 (hardware log2), with
//   log2 U_b <= log2 nf + max_k (log2 c_bk - alpha_bk log2(e) g(x_b)),
//   log2 L   >= max(max_k (log2 c_ak - alpha_ak log2(e) g(rhi)), log2 floors),
// g(x) = x (Slater) or x^2 (Gauss), and 2 units of margin for the FP32
// logarithms and products; distances carry 1e-6 relative margins, which also
// cover the rounding of the kernel's FP32 distances.
#define HP_TAIL_MIN_ATOMS 32  // fewer candidates: the bound costs more than it saves
#define HP_TAIL_LOG2_EPS64 (-64.0f)
#define HP_TAIL_LOG2_EPS32 (-40.0f)

template <int FK>
__device__ __forceinline__ float tail_log2_term(double c, double al, float x) {
  const float g = FK == HP_FN_GAUSS ? x * x : x;
  return __builtin_amdgcn_logf((float)c) - (float)(al * 1.4426950408889634) * g;
}

// log2 of the lower bound of the promolecule on the shell (without margin)
template <int FK>
__device__ __forceinline__ float tail_log2_own(const DevFn &fn, int a, float rhi, float log2_floors) {
  const int f0 = fn.fn_off[a], nf = fn.fn_off[a + 1] - f0;
  const float x = rhi * (1.0f + 1e-6f);
  float m = log2_floors;
  for (int k = 0; k < nf; ++k) m = fmaxf(m, tail_log2_term<FK>(fn.coef[f0 + k], fn.alpha[f0 + k], x));
  return m;
}

// The bound log2 U_b for one candidate (its values are consumed as they
// arrive, so few registers stay live beside the points').
template <int FK>
__device__ __forceinline__ float tail_log2_bound(const DevMol &m, const DevFn &fn, int b, double ax, double ay,
                                                 double az, float rhi) {
  const float dx = (float)(m.cx[b] - ax), dy = (float)(m.cy[b] - ay), dz = (float)(m.cz[b] - az);
  const float d = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
  const float x = fmaxf(0.0f, d * (1.0f - 1e-6f) - rhi * (1.0f + 1e-6f));
  const int f0 = fn.fn_off[b], nf = fn.fn_off[b + 1] - f0;
  float u = -__builtin_inff();
  for (int k = 0; k < nf; ++k) u = fmaxf(u, tail_log2_term<FK>(fn.coef[f0 + k], fn.alpha[f0 + k], x));
  return u + __builtin_amdgcn_logf((float)nf);
}

// The FP32 tier of the FP64 pass (no cut-off radius).  An atom whose bound
// is below 2^-42 L adds less than 2^-42 of the promolecule at every point of
// the workgroup.  Evaluated in FP32 (offsets from the shell's atom formed in
// FP64 and rounded once, v_sqrt_f32, v_exp_f32 with the exponent scaled by
// log2 e), a term that FP32 represents carries a relative error below 3e-4:
// the distance is within 1.8e-7 (r + 2 r_max) of the FP64 one, and the
// exponent alpha log2(e) r is below 136 for such a term, so with alpha up to
// 40 per bohr and r_max up to 20 bohr the exponent moves by less than 4.4e-4.
// The promolecule then moves by less than 3e-4 x 2^-42 < 2^-53 relative,
// under one FP64 rounding of it.  Terms below the FP32 range (1.2e-38) are
// lost; the tier is used only where L >= 2^-60, so that their sum over a few
// hundred atoms stays below 2^-53 L as well.  Floors are added in FP64 as for
// every atom.
#define HP_TAIL_LOG2_TIER32 (-42.0f)
#define HP_TAIL_LOG2_TIER32_MIN_L (-60.0f)

template <int FK, int PT, int NP, int KM>
__device__ __forceinline__ int weights_tile_f32(const float *__restrict__ rec, int tn, const double (&px)[PT],
                                                const double (&py)[PT], const double (&pz)[PT], double ax, double ay,
                                                double az, double (&promol)[PT]) {
  constexpr int RS = weights_record<float>(KM);
  // the points' offsets from the shell's atom, formed in FP64 and rounded once
  float qx[NP], qy[NP], qz[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    qx[j] = (float)(px[j] - ax);
    qy[j] = (float)(py[j] - ay);
    qz[j] = (float)(pz[j] - az);
  }
  int nfn = 0;
  for (int i = 0; i < tn; ++i) {
    const float *q = rec + (size_t)i * RS;
    float v[RS];
#pragma unroll
    for (int u = 0; u < RS; ++u) v[u] = q[u];
    const int nf = __builtin_amdgcn_readfirstlane((int)v[3]);
    nfn += nf;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const float dx = qx[j] - v[0], dy = qy[j] - v[1], dz = qz[j] - v[2];
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float arg = FK == HP_FN_GAUSS ? r2 : __builtin_amdgcn_sqrtf(r2);
      float pb = 0.0f;
#pragma unroll
      for (int k = 0; k < KM; ++k)
        if (k < nf) pb += v[4 + k] * __builtin_amdgcn_exp2f(-v[4 + KM + k] * arg);
      promol[j] += (double)pb;
      promol[j] += 1e-100;
    }
  }
  return nfn;
}

// Defect correction of the FP32 pass (corr_mode):
//   HP_CORR_NONE      rhobar = this pass's average
//   HP_CORR_APPLY     rhobar = FP32 average + corr (the correction of the last refresh)
//   HP_CORR_CALIBRATE corr = rhobar - FP32 average, where rhobar was just written by
//                     the FP64 pass at the same parameters (rhobar, entropy kept)
//   HP_CORR_SLOT      batched: per molecule, CALIBRATE on its refresh iterations,
//                     APPLY otherwise (FP32 launches); FP64 launches also run the
//                     FP32-stage molecules that refresh in this iteration.
#define HP_CORR_NONE 0
#define HP_CORR_APPLY 1
#define HP_CORR_CALIBRATE 2
#define HP_CORR_SLOT -1

template <typename R, int FK, bool GT2, int PT, int KM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(R) == sizeof(double) ? (KM <= 2 ? 5 : (KM <= 4 ? 4 : 1)) : (KM <= 2 ? 6 : (KM <= 4 ? 5 : 1))))) void k_weights(const int *__restrict__ done, DevMol m, DevFn fn,
                                                 double *__restrict__ rhobar, double *__restrict__ ent_part,
                                                 double *__restrict__ promol_out, double dc, int tile,
                                                 double *__restrict__ corr, int corr_mode) {
  // atom records of a staging round; at the end the points' terms
  extern __shared__ __align__(16) double smem_d[];
  __shared__ int s_own;  // list position of the shell's own atom
  __shared__ double s_t64[64];  // 2^(j/64) for the FP64 exponential
  __shared__ float s_rhi[8];    // per wave: largest point distance from the shell's atom
  __shared__ int s_cnt[8];      // per wave: candidates kept in this staging round
  __shared__ int s_cnt32[8];    // per wave: of those, in the FP32 tier
  __shared__ int s_own_tile, s_own_pos;  // tail path: staging round and slot of the own atom
  if (done && *done) return;
  constexpr int RS = weights_record<R>(KM);
  R *s_rec = reinterpret_cast<R *>(smem_d);
  int *s_idx = reinterpret_cast<int *>(s_rec + (size_t)tile * RS);
  float *s_rec32 = reinterpret_cast<float *>(s_idx + tile);  // FP64 pass: the FP32 tier's records
  constexpr int RS32 = weights_record<float>(KM);

  int s = m.grp_first ? m.grp_first[blockIdx.x] : (int)blockIdx.x;  // first radial shell of the group
  int a_begin = 0, a_end = m.natom;  // atoms of this shell's molecule
  int mode = corr_mode;
  int full = 0;
  int wgroup = m.wgroup;
  if (m.work) {
    const int2 wk = m.work[blockIdx.x];
    const HpSlot &sl = m.slots[wk.y];
    if (sl.phase) return;  // its inner solves of this outer iteration are still running
    full = sl.full_sum;
    wgroup = sl.wgroup;
    // a molecule takes part in the launches of its current stage, and an
    // FP32-stage molecule also in the FP64 launch of a refresh iteration
    if (sizeof(R) == sizeof(double)) {
      if (!(sl.run == HP_RUN_FP64 || (sl.run == HP_RUN_FP32 && sl.refresh))) return;
      mode = HP_CORR_NONE;
    } else {
      if (sl.run != HP_RUN_FP32) return;
      if (mode == HP_CORR_SLOT) mode = sl.corr_period > 0 ? (sl.refresh ? HP_CORR_CALIBRATE : HP_CORR_APPLY) : HP_CORR_NONE;
    }
    s = wk.x;
    a_begin = sl.a0;
    a_end = sl.a0 + sl.na;
  }
  const int a = m.shell_atom[s];
  // the group: up to wgroup consecutive radial shells of atom a, their points
  // contiguous and spread over the workgroup's point slots
  const int gn = min(max(wgroup, 1), m.atom_shell_off[a + 1] - s);
  const int64_t p0 = m.shell_off[s], p1 = m.shell_off[s + gn];
  int nb_n = a_end - a_begin;
  const int *nb = nullptr;
  const bool cull = fn.rc > 0.0;
  if (cull) {
    // each shell's candidates are a prefix of the atom's list sorted by
    // distance: the group takes the longest; the per-point cutoff test
    // drops the extra atoms on its inner shells' points
    nb = fn.nb_list + (fn.nb_off ? fn.nb_off[a] : (int64_t)a * m.natom);
    nb_n = 0;
    for (int g = 0; g < gn; ++g) nb_n = max(nb_n, fn.shell_nb[s + g]);
  }
  const R rc = (R)fn.rc;
  const double ax = m.cx[a], ay = m.cy[a], az = m.cz[a];
  int64_t pidx[PT];
  double px[PT], py[PT], pz[PT];
  R qx[PT], qy[PT], qz[PT];
  R promol[PT], own[PT];
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    pidx[j] = p0 + threadIdx.x + (int64_t)j * blockDim.x;
    const bool ok = pidx[j] < p1;
    px[j] = ok ? m.px[pidx[j]] : ax;
    py[j] = ok ? m.py[pidx[j]] : ay;
    pz[j] = ok ? m.pz[pidx[j]] : az;
    qx[j] = (R)(px[j] - ax);
    qy[j] = (R)(py[j] - ay);
    qz[j] = (R)(pz[j] - az);
    promol[j] = (R)0;
    own[j] = (R)0;
  }
  const bool two = p0 + blockDim.x < p1;  // second point slot used by any thread
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (threadIdx.x == 0) {
    s_own = -1;
    s_own_tile = -1;
  }
  if (sizeof(R) == sizeof(double) && threadIdx.x < 64) s_t64[threadIdx.x] = hp_exp2_64[threadIdx.x];
  // bounded tails: the shell's largest point distance (published by the
  // first staging barrier).  The host enables them only for coefficients that
  // stay non-negative (MBIS, GISA, from non-negative initial parameters).
  const bool tail_req = fn.tail_cull && !full && FK != HP_FN_GENERAL && p0 < p1 && nb_n >= HP_TAIL_MIN_ATOMS;
  if (tail_req) {
    float rhi = -__builtin_inff();
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      if (pidx[j] < p1) {
        const float dx = (float)(px[j] - ax), dy = (float)(py[j] - ay), dz = (float)(pz[j] - az);
        rhi = fmaxf(rhi, __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz));
      }
    }
    rhi = (float)-wave_min(-(double)rhi);
    if (lane == 0) s_rhi[wv] = rhi;
  }
  const bool tail = tail_req;
  float tail_rhi = 0.0f, tail_log2_thr = 0.0f, tail_log2_thr32 = 0.0f;
  bool tier32 = false;  // FP64 pass: the FP32 tier in use for this workgroup
  int n_cut = 0;  // candidates skipped by the tail bound
  int n_fn = 0;  // functions evaluated on every point of this wavefront (uniform in it)
  int n_fn32 = 0;  // of the FP32 tier
  // the tail path stages one candidate per thread and round
  const int step = tail_req ? min(tile, (int)blockDim.x) : tile;
  for (int t0 = 0, round = 0; t0 < nb_n; t0 += step, ++round) {
    const int tn = min(step, nb_n - t0);
    __syncthreads();
    int tk = tn, tk32 = 0, own_i;
    if (!tail) {
      for (int i = threadIdx.x; i < tn; i += blockDim.x) {
        const int b = nb ? nb[t0 + i] : a_begin + t0 + i;
        stage_record<R, KM>(s_rec + (size_t)i * RS, s_idx + i, b, m, fn, ax, ay, az);
        if (b == a) s_own = t0 + i;
      }
      __syncthreads();
      own_i = __builtin_amdgcn_readfirstlane(s_own - t0);  // outside [0, tn) when not in this tile
    } else {
      // keep flags, then the kept candidates packed in list order
      const int i = threadIdx.x;
      int b = -1;
      bool keep = false;
      if (i < tn) b = nb ? nb[t0 + i] : a_begin + t0 + i;
      if (t0 == 0) {
        // the threshold (its loads overlap the candidates')
        float rhi = -__builtin_inff();
        for (int w = 0; w < nw; ++w) rhi = fmaxf(rhi, s_rhi[w]);
        tail_rhi = rhi;
        // log2 of the floors' sum (FP64: n x 1e-100 = 2^(log2 n - 332.19); FP32 adds none)
        const float log2_floors = sizeof(R) == sizeof(double)
                                      ? __builtin_amdgcn_logf((float)(a_end - a_begin)) - 332.19281f
                                      : -__builtin_inff();
        const float log2_eps = sizeof(R) == sizeof(double) ? HP_TAIL_LOG2_EPS64 : HP_TAIL_LOG2_EPS32;
        const float log2_own = tail_log2_own<FK>(fn, a, rhi, log2_floors);
        tail_log2_thr = log2_eps - 2.0f + log2_own;
        tail_log2_thr32 = HP_TAIL_LOG2_TIER32 - 2.0f + log2_own;
        tier32 = sizeof(R) == sizeof(double) && !cull && log2_own >= HP_TAIL_LOG2_TIER32_MIN_L;
      }
      bool keep32 = false;
      if (i < tn) {
        if (b == a) {
          keep = true;
        } else {
          const float ub = tail_log2_bound<FK>(m, fn, b, ax, ay, az, tail_rhi);
          keep = !(ub < tail_log2_thr);
          keep32 = keep && tier32 && ub < tail_log2_thr32;
        }
      }
      const bool keep64 = keep && !keep32;
      const unsigned long long bal = __ballot(keep64), bal32 = __ballot(keep32);
      if (lane == 0) {
        s_cnt[wv] = __popcll(bal);
        s_cnt32[wv] = __popcll(bal32);
      }
      __syncthreads();
      int base = 0, base32 = 0;
      tk = 0;
      for (int w = 0; w < nw; ++w) {
        const int c = s_cnt[w], c32 = s_cnt32[w];
        base += w < wv ? c : 0;
        base32 += w < wv ? c32 : 0;
        tk += c;
        tk32 += c32;
      }
      if (keep64) {
        const int pos =
            base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
        stage_record<R, KM>(s_rec + (size_t)pos * RS, s_idx + pos, b, m, fn, ax, ay, az);
        if (b == a) {
          s_own_tile = round;
          s_own_pos = pos;
        }
      }
      if (keep32) {
        const int pos = base32 + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal32 >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)bal32, 0u));
        int unused;
        stage_record<float, KM>(s_rec32 + (size_t)pos * RS32, &unused, b, m, fn, ax, ay, az);
      }
      __syncthreads();
      // uniform across the block: scalar registers keep the atom loop uniform
      tk = __builtin_amdgcn_readfirstlane(tk);
      tk32 = __builtin_amdgcn_readfirstlane(tk32);
      own_i = __builtin_amdgcn_readfirstlane(s_own_tile == round ? s_own_pos : -1);
      n_cut += tn - tk - tk32;
    }
    // every point slot in use, or only the first (the shell fits one pass of the block)
    if (cull) {
      n_fn += two ? weights_tile<R, FK, PT, PT, KM, true>(s_rec, s_idx, s_t64, tk, own_i, fn, rc, qx, qy, qz, px, py,
                                                          pz, promol, own)
                  : weights_tile<R, FK, PT, 1, KM, true>(s_rec, s_idx, s_t64, tk, own_i, fn, rc, qx, qy, qz, px, py,
                                                         pz, promol, own);
    } else {
      n_fn += two ? weights_tile<R, FK, PT, PT, KM, false>(s_rec, s_idx, s_t64, tk, own_i, fn, rc, qx, qy, qz, px, py,
                                                           pz, promol, own)
                  : weights_tile<R, FK, PT, 1, KM, false>(s_rec, s_idx, s_t64, tk, own_i, fn, rc, qx, qy, qz, px, py,
                                                          pz, promol, own);
      if constexpr (sizeof(R) == sizeof(double)) {
        if (tk32 > 0)
          n_fn32 += two ? weights_tile_f32<FK, PT, PT, KM>(s_rec32, tk32, px, py, pz, ax, ay, az, promol)
                        : weights_tile_f32<FK, PT, 1, KM>(s_rec32, tk32, px, py, pz, ax, ay, az, promol);
      }
    }
  }
  // per point: its spherical-average and entropy terms, staged in LDS over
  // the atom records (no longer read after this barrier)
  const int64_t npt = p1 - p0;
  double *s_avg = smem_d, *s_ent = smem_d + npt;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    if (pidx[j] >= p1) continue;
    const int64_t p = pidx[j];
    R pm_r = promol[j];
    // culled and skipped atoms still contribute their 1e-100 floor
    const int n_floor = (a_end - a_begin - nb_n) + n_cut;
    if (n_floor) pm_r += (R)n_floor * hp_floor_add<R>();
    const double pm = (double)pm_r;
    const double rho = m.rho[p];
    double f_avg = 0.0, f_ent = 0.0;
    if (!GT2) {
      const double w = aim_weight(own[j], pm_r);
      const double f = w * rho;
      f_avg = m.wsph[p] * f;
    } else {
      promol_out[p] = pm;
    }
    // entropy integrand (a diagnostic): the log in R, the sum in FP64
    if (!(pm < dc || rho < dc)) {
      const double lg =
          sizeof(R) == sizeof(double) ? log(rho / pm) : (double)__logf((float)rho / (float)pm_r);
      f_ent = m.wmol[p] * rho * lg;
    }
    const int q = (int)threadIdx.x + j * (int)blockDim.x;  // p - p0
    s_avg[q] = f_avg;
    s_ent[q] = f_ent;
  }
  __syncthreads();
  // one wave per radial shell of the group: lanes add its points in a fixed
  // stride order, then a wave all-reduce
  for (int g = __builtin_amdgcn_readfirstlane(wv); g < gn; g += nw) {
    const int q0 = (int)(m.shell_off[s + g] - p0), q1 = (int)(m.shell_off[s + g + 1] - p0);
    double v[2] = {0.0, 0.0};
    for (int q = q0 + (int)(threadIdx.x & 63); q < q1; q += 64) {
      v[0] += s_avg[q];
      v[1] += s_ent[q];
    }
    wave_allsum_n<2>(v);
    if (lane == 0) {
      const int sg = s + g;
      if (GT2) {
        ent_part[sg] = v[1];
      } else if (mode == HP_CORR_CALIBRATE) {
        corr[sg] = rhobar[sg] - v[0];
      } else {
        rhobar[sg] = mode == HP_CORR_APPLY ? v[0] + corr[sg] : v[0];
        ent_part[sg] = v[1];
      }
    }
  }
  if (fn.evals_done) {
    // per wavefront: its evaluated functions times its points
    unsigned long long npts_w = 0;
#pragma unroll
    for (int j = 0; j < PT; ++j) npts_w += (unsigned long long)__popcll(__ballot(pidx[j] < p1));
    if (lane == 0 && npts_w) {
      const int lane_slot = (int)((blockIdx.x * 8u + wv) & (HP_EVAL_LANES - 1));
      const int slot = (sizeof(R) == sizeof(double) ? HP_EVAL_LANES : 0) + lane_slot;
      atomicAdd(fn.evals_done + (size_t)slot * HP_EVAL_STRIDE, (unsigned long long)n_fn * npts_w);
      // the FP32 tier's exponentials count with the FP32 ones
      if (n_fn32) atomicAdd(fn.evals_done + (size_t)lane_slot * HP_EVAL_STRIDE, (unsigned long long)n_fn32 * npts_w);
    }
  }
}

// Promolecule (stored) or AIM weights of one atom over a point range.
template <int FK>
__global__ void k_point_weights(DevMol m, DevFn fn, int atom, int64_t p0, int64_t n, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = p0 + i;
  double promol, own;
  // culling lists are per owning atom and shell; evaluate without them here
  DevFn f2 = fn;
  f2.nb_list = nullptr;
  const int a = atom < 0 ? 0 : atom;
  promol_at<double, FK>(m, f2, a, m.natom, nullptr, m.px[p], m.py[p], m.pz[p], promol, own);
  if (atom < 0) {
    out[i] = promol;
  } else {
    out[i] = aim_weight(own, promol);
  }
}

// ---------------------------------------------------------------------------
// per-atom radial data shared by the inner solvers
// ---------------------------------------------------------------------------
template <int PPL>
struct RadialLane {
  double r[PPL], w[PPL], wr[PPL], rho[PPL];
};

template <int PPL>
__device__ __forceinline__ void load_radial(const DevMol &m, const double *rhobar, int s0, int R, int lane,
                                            RadialLane<PPL> &d) {
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int i = lane + 64 * j;
    if (i < R) {
      const double r = m.shell_r[s0 + i];
      d.r[j] = r;
      d.wr[j] = m.shell_w[s0 + i];
      d.w[j] = 4.0 * HP_PI * (r * r) * d.wr[j];
      d.rho[j] = rhobar[s0 + i];
    } else {
      d.r[j] = 0.0;
      d.wr[j] = 0.0;
      d.w[j] = 0.0;
      d.rho[j] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// K2: MBIS inner fixed point (mbis.py:81-163), one workgroup per atom with one
// radial point per lane (ceil(R/64) waves).  Each iteration reduces 2K+1
// values: an all-reduce inside every wave (DPP and lane swaps), then the wave
// totals through LDS in wave order, so all lanes of all waves hold bitwise
// identical sums and take the same convergence branch.  The LDS slots alternate with the iteration
// parity, which needs one barrier per iteration.
// KM: compile-time bound on shells per atom (loops unrolled, guarded by K).
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_allsum(double *v, double *lds, int parity) {
  // lds: 2 x 16 waves x NV doubles
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  wave_allsum_n<NV>(v);
  double *slot = lds + (size_t)parity * 16 * NV;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) slot[wv * NV + i] = v[i];
  }
  __syncthreads();
  double t[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) t[i] = 0.0;
  for (int w = 0; w < nw; ++w) {
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] += slot[w * NV + i];
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = t[i];
}

#define HP_INV_8PI 0.039788735772973836  // 1 / (8 pi), rounded to the nearest double

// Per-block context of one atom's MBIS solve.
#define HP_MBIS_PPL_MAX 4  // one wavefront per atom up to 256 radial points

struct MbisCtx {
  const DevMol *m;
  double *par;
  const double *old_par;
  const double *s_t64;
  double *lds;
  int *inner_niter, *status, *ist_it;
  double *charges, *change_part, *ist_oldpro;
  double thr, pop;
  // this lane's radial points i + 64 j, j < PPL (the solver's template)
  double r[HP_MBIS_PPL_MAX], wr[HP_MBIS_PPL_MAX], w[HP_MBIS_PPL_MAX], rho[HP_MBIS_PPL_MAX];
  bool valid[HP_MBIS_PPL_MAX];
  int a, f0, s0, i, maxit, budget;
  bool resume;
};

// All-reduce of the solver: across the workgroup (one radial point per lane,
// several wavefronts) or inside the one wavefront (two points per lane).
template <int NV, int PPL>
__device__ __forceinline__ void solver_allsum(double *v, double *lds, int parity) {
  if constexpr (PPL == 1)
    block_allsum<NV>(v, lds, parity);
  else if constexpr (NV == 1)
    wave_allsum_n<1>(v);
  else
    wave_allsum_packed<NV>(v);
}

// The solve for an atom of exactly KC shells (no per-shell guards in the
// loop; 2 KC + 1 sums per iteration), PPL radial points per lane.
template <int KC, int PPL>
__device__ __forceinline__ void mbis_solve(const MbisCtx &c, int slot) {
  constexpr int NV = 2 * KC + 1;
  double Np[KC], Sp[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    Np[k] = c.par[2 * (c.f0 + k)];
    Sp[k] = c.par[2 * (c.f0 + k) + 1];
  }
  // a solve resumed from an earlier launch continues from its saved state:
  // the iterate (in par), the previous pro-atom and the iterations done
  const int it0 = c.resume ? c.ist_it[c.a] : 0;
  double oldpro[PPL];
#pragma unroll
  for (int j = 0; j < PPL; ++j) oldpro[j] = (c.resume && c.valid[j]) ? c.ist_oldpro[c.s0 + c.i + 64 * j] : 0.0;
  bool have_old = it0 > 0, conv = false;
  int it = 0;
  const int it_end = (c.budget > 0 && it0 + c.budget < c.maxit) ? it0 + c.budget : c.maxit;
  const double dc = 1e-15;  // opt_mbis_propars default density_cutoff (mbis.py:82)
  // The stop test sqrt(S) < thr (S = the summed squared change) decided
  // from S where the bounds thr^2 (1 -/+ 2^-50) settle it (they cover the
  // roundings of thr^2 and of the square root), so the square root runs only
  // in that narrow band: the same decision, off the loop's critical path.
  const double thr2 = c.thr * c.thr;
  // (a threshold <= 0 never stops: no bound passes)
  const double thr2_lo = c.thr > 0.0 ? thr2 * (1.0 - 0x1p-50) : -1.0;
  const double thr2_hi = c.thr > 0.0 ? thr2 * (1.0 + 0x1p-50) : -1.0;
  for (it = it0; it < it_end; ++it) {
    double red[NV];
    double pro[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const double r = c.r[j], w = c.w[j], rho = c.rho[j];
      double t[KC];
      pro[j] = 0.0;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        // N S^3 / (8 pi), once per shell; the constant division as a product
        // with its rounded reciprocal (within one rounding of the quotient)
        const double coef = Np[k] * (Sp[k] * Sp[k] * Sp[k]) * HP_INV_8PI;
        t[k] = coef * exp_tab(-Sp[k] * r, c.s_t64);
        pro[j] += t[k];
      }
      const bool sick = (rho < dc) || (pro[j] < dc);
      const double ratio = sick ? 0.0 : rho / pro[j];
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const double wt = w * (t[k] * ratio);
        red[2 * k] = j ? red[2 * k] + wt : wt;
        red[2 * k + 1] = j ? red[2 * k + 1] + wt * r : wt * r;
      }
      const double e = oldpro[j] - pro[j];
      red[2 * KC] = j ? red[2 * KC] + w * e * e : w * e * e;
    }
    solver_allsum<NV, PPL>(red, c.lds, slot);
    slot ^= 1;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      Np[k] = red[2 * k];
      Sp[k] = 3.0 * red[2 * k] / red[2 * k + 1];
    }
    if (have_old) {
      const double S = red[2 * KC];
      // a squared change is 0 or far above 2^-767: the square root needs no rescaling
      const bool stop = S < thr2_lo || (S <= thr2_hi && fast_sqrt<double>(S) < c.thr);
      if (stop) {
        conv = true;
        break;
      }
    }
#pragma unroll
    for (int j = 0; j < PPL; ++j) oldpro[j] = pro[j];
    have_old = true;
  }
  if (!conv && it < c.maxit) {
    // iteration budget of this launch spent: keep the state for the next one
#pragma unroll
    for (int j = 0; j < PPL; ++j)
      if (c.valid[j]) c.ist_oldpro[c.s0 + c.i + 64 * j] = oldpro[j];
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        c.par[2 * (c.f0 + k)] = Np[k];
        c.par[2 * (c.f0 + k) + 1] = Sp[k];
      }
      c.ist_it[c.a] = it;
    }
    return;
  }
  if (c.ist_it && threadIdx.x == 0) c.ist_it[c.a] = -1;
  // change of the outer loop: new vs old pro-atom on the radial grid
  double msd = 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const double r = c.r[j];
    double y1 = 0.0, y2 = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const double N1 = Np[k], S1 = Sp[k];
      const double N2 = c.old_par[2 * (c.f0 + k)], S2 = c.old_par[2 * (c.f0 + k) + 1];
      y1 += N1 * (S1 * S1 * S1) * HP_INV_8PI * exp_tab(-S1 * r, c.s_t64);
      y2 += N2 * (S2 * S2 * S2) * HP_INV_8PI * exp_tab(-S2 * r, c.s_t64);
    }
    const double dy = y1 - y2;
    msd += c.wr[j] * (4.0 * HP_PI * (r * r)) * dy * dy;
  }
  solver_allsum<1, PPL>(&msd, c.lds, slot);
  if (threadIdx.x == 0) {
    int st = 0;
    double sumN = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) sumN += Np[k];
    if (conv) {
      // np.isclose(pop, sum, atol=1e-4): |pop - sum| <= 1e-4 + 1e-5 |sum|
      if (!(fabs(c.pop - sumN) <= 1e-4 + 1e-5 * fabs(sumN))) st |= HP_ST_POP;
    } else {
      st |= HP_ST_NOCONV;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      c.par[2 * (c.f0 + k)] = Np[k];
      c.par[2 * (c.f0 + k) + 1] = Sp[k];
    }
    c.inner_niter[c.a] = conv ? it + 1 : c.maxit;
    c.status[c.a] = st;
    c.charges[c.a] = c.m->pseudo[c.a] - c.pop;
    c.change_part[c.a] = msd;
  }
}

// The solve for the atom's shell count K in [1, KM] (uniform per block).
template <int KM, int KC, int PPL>
__device__ __forceinline__ void mbis_dispatch(int K, const MbisCtx &c, int slot) {
  if constexpr (KC <= KM) {
    if (K == KC) {
      mbis_solve<KC, PPL>(c, slot);
      return;
    }
    mbis_dispatch<KM, KC + 1, PPL>(K, c, slot);
  }
}

// PPL = 1: one workgroup per atom, one radial point per lane (ceil(R/64)
// wavefronts).  PPL = 2, 3, 4: one wavefront per atom, radial points i,
// i + 64, ... on lane i (R <= 64 PPL): its sums need no LDS exchange and no
// barrier.
template <int KM, int PPL>
__global__ __launch_bounds__(512) void k_mbis_inner(const int *__restrict__ done, DevMol m,
                                                    const double *__restrict__ rhobar, double *__restrict__ par,
                                                    const double *__restrict__ old_par, const int *__restrict__ fn_off,
                                                    double thr, int maxit, int *__restrict__ inner_niter,
                                                    int *__restrict__ status, double *__restrict__ charges,
                                                    double *__restrict__ change_part, int budget,
                                                    int *__restrict__ ist_it, double *__restrict__ ist_oldpro) {
  constexpr int NV = 2 * KM + 1;
  __shared__ double lds[2 * 16 * NV];
  __shared__ double s_t64[64];  // 2^(j/64) for the exponential
  if (done && *done) return;
  // a dependent chain of short steps: its waves issue first on their SIMDs,
  // beside the throughput-bound weights pass of other molecules
  __builtin_amdgcn_s_setprio(3);
  if (threadIdx.x < 64) s_t64[threadIdx.x] = hp_exp2_64[threadIdx.x];
  int a = blockIdx.x;
  bool resume = false;
  if (m.work) {
    const int2 wk = m.work[blockIdx.x];
    const HpSlot &sl = m.slots[wk.y];
    if (sl.run == HP_RUN_IDLE) return;
    a = wk.x;
    if (sl.phase) {
      if (ist_it[a] < 0) return;  // this atom's solve ended in an earlier launch
      resume = true;
    }
  }
  __syncthreads();  // the table (uniform exits above: the whole block returns)
  const int s0 = m.atom_shell_off[a], R = m.atom_shell_off[a + 1] - s0;
  const int f0 = fn_off[a], K = __builtin_amdgcn_readfirstlane(fn_off[a + 1] - f0);
  const int i = threadIdx.x;
  MbisCtx c;
  c.m = &m;
  c.par = par;
  c.old_par = old_par;
  c.s_t64 = s_t64;
  c.lds = lds;
  c.inner_niter = inner_niter;
  c.status = status;
  c.ist_it = ist_it;
  c.charges = charges;
  c.change_part = change_part;
  c.ist_oldpro = ist_oldpro;
  c.thr = thr;
  double pv = 0.0;
#pragma unroll
  for (int j = 0; j < HP_MBIS_PPL_MAX; ++j) {
    const int q = i + 64 * j;
    const bool valid = j < PPL && q < R;
    c.valid[j] = valid;
    c.r[j] = valid ? m.shell_r[s0 + q] : 0.0;
    c.wr[j] = valid ? m.shell_w[s0 + q] : 0.0;
    c.w[j] = 4.0 * HP_PI * (c.r[j] * c.r[j]) * c.wr[j];
    c.rho[j] = valid ? rhobar[s0 + q] : 0.0;
    if (j < PPL) pv = j ? pv + c.w[j] * c.rho[j] : c.w[j] * c.rho[j];
  }
  c.a = a;
  c.f0 = f0;
  c.s0 = s0;
  c.i = i;
  c.maxit = maxit;
  c.budget = budget;
  c.resume = resume;
  int slot = 0;  // LDS slot of the next block reduction (alternates)
  solver_allsum<1, PPL>(&pv, lds, slot);
  slot ^= 1;
  c.pop = pv;
  mbis_dispatch<KM, 1, PPL>(K, c, slot);
}

// ---------------------------------------------------------------------------
// NLIS / GMBIS inner fixed point (nlis.py:99-195): shells N g(n, S)
// exp(-S r^n), g(n, S) = n S^(3/n) / (4 pi Gamma(3/n)), parameters (N, S, n)
// per shell with n fixed.  Per iteration N <- sum w t ratio N and
// S <- 3 / (n sum w t ratio r^n) (1e-5 when that moment is 0 to 1e-8), t the
// unit-population shell.  One workgroup per atom, one radial point per lane;
// gnorm[f] = n / (4 pi Gamma(3/n)) per function (fixed with n).
// ---------------------------------------------------------------------------
template <int KC>
__device__ __forceinline__ void nlis_solve(const DevMol &m, int a, int f0, int s0, int R, double *par,
                                           const double *old_par, const double *gnorm, const double *rhobar,
                                           double thr, int maxit, int *inner_niter, int *status, double *charges,
                                           double *change_part, double *lds) {
  constexpr int NV = 2 * KC + 1;
  const int i = threadIdx.x;
  const bool valid = i < R;
  const double r = valid ? m.shell_r[s0 + i] : 0.0;
  const double wr = valid ? m.shell_w[s0 + i] : 0.0;
  const double w = 4.0 * HP_PI * (r * r) * wr;
  const double rho = valid ? rhobar[s0 + i] : 0.0;
  double Np[KC], Sp[KC], np_[KC], rn[KC], g[KC], e3[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    Np[k] = par[3 * (f0 + k)];
    Sp[k] = par[3 * (f0 + k) + 1];
    np_[k] = par[3 * (f0 + k) + 2];
    rn[k] = pow(r, np_[k]);
    g[k] = gnorm[f0 + k];
    e3[k] = 3.0 / np_[k];
  }
  int slot = 0;
  double pv = w * rho;
  block_allsum<1>(&pv, lds, slot);
  slot ^= 1;
  const double pop = pv;
  const double dc = 1e-15;  // opt_nlis_propars default density_cutoff (nlis.py:100)
  double oldpro = 0.0;
  bool have_old = false, conv = false;
  int it = 0;
  for (it = 0; it < maxit; ++it) {
    double red[NV];
    double t[KC];
    double pro = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      t[k] = g[k] * pow(Sp[k], e3[k]) * exp(-Sp[k] * rn[k]);
      pro += t[k] * Np[k];
    }
    const bool sick = (rho < dc) || (pro < dc);
    const double ratio = sick ? 0.0 : rho / pro;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const double tr = t[k] * ratio;
      red[2 * k] = w * (tr * Np[k]);
      red[2 * k + 1] = w * tr * rn[k];
    }
    const double e = oldpro - pro;
    red[2 * KC] = w * e * e;
    block_allsum<NV>(red, lds, slot);
    slot ^= 1;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      Np[k] = red[2 * k];
      const double m1 = red[2 * k + 1];
      // np.isclose(m1, 0.0): |m1| <= 1e-8
      Sp[k] = fabs(m1) <= 1e-8 ? 1e-5 : 3.0 / (m1 * np_[k]);
    }
    if (have_old && sqrt(red[2 * KC]) < thr) {
      conv = true;
      break;
    }
    oldpro = pro;
    have_old = true;
  }
  // change of the outer loop: new vs old pro-atom on the radial grid
  double y1 = 0.0, y2 = 0.0;
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const double S2 = old_par[3 * (f0 + k) + 1];
    y1 += Np[k] * g[k] * pow(Sp[k], e3[k]) * exp(-Sp[k] * rn[k]);
    y2 += old_par[3 * (f0 + k)] * g[k] * pow(S2, e3[k]) * exp(-S2 * rn[k]);
  }
  const double dy = y1 - y2;
  double msd = wr * (4.0 * HP_PI * (r * r)) * dy * dy;
  block_allsum<1>(&msd, lds, slot);
  if (threadIdx.x == 0) {
    int st = 0;
    double sumN = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) sumN += Np[k];
    if (conv) {
      if (!(fabs(pop - sumN) <= 1e-4 + 1e-5 * fabs(sumN))) st |= HP_ST_POP;
    } else {
      st |= HP_ST_NOCONV;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      par[3 * (f0 + k)] = Np[k];
      par[3 * (f0 + k) + 1] = Sp[k];
    }
    inner_niter[a] = conv ? it + 1 : maxit;
    status[a] = st;
    charges[a] = m.pseudo[a] - pop;
    change_part[a] = msd;
  }
}

template <int KC, int KM>
__device__ __forceinline__ void nlis_dispatch(int K, const DevMol &m, int a, int f0, int s0, int R, double *par,
                                              const double *old_par, const double *gnorm, const double *rhobar,
                                              double thr, int maxit, int *inner_niter, int *status, double *charges,
                                              double *change_part, double *lds) {
  if constexpr (KC <= KM) {
    if (K == KC) {
      nlis_solve<KC>(m, a, f0, s0, R, par, old_par, gnorm, rhobar, thr, maxit, inner_niter, status, charges,
                     change_part, lds);
      return;
    }
    nlis_dispatch<KC + 1, KM>(K, m, a, f0, s0, R, par, old_par, gnorm, rhobar, thr, maxit, inner_niter, status,
                              charges, change_part, lds);
  }
}

#define HP_NLIS_KM 8
static __global__ __launch_bounds__(512) void k_nlis_inner(const int *__restrict__ done, DevMol m,
                                                    const double *__restrict__ rhobar, double *__restrict__ par,
                                                    const double *__restrict__ old_par, const int *__restrict__ fn_off,
                                                    const double *__restrict__ gnorm, double thr, int maxit,
                                                    int *__restrict__ inner_niter, int *__restrict__ status,
                                                    double *__restrict__ charges, double *__restrict__ change_part) {
  __shared__ double lds[2 * 16 * (2 * HP_NLIS_KM + 1)];
  if (done && *done) return;
  const int a = blockIdx.x;
  const int s0 = m.atom_shell_off[a], R = m.atom_shell_off[a + 1] - s0;
  const int f0 = fn_off[a], K = __builtin_amdgcn_readfirstlane(fn_off[a + 1] - f0);
  nlis_dispatch<1, HP_NLIS_KM>(K, m, a, f0, s0, R, par, old_par, gnorm, rhobar, thr, maxit, inner_niter, status,
                               charges, change_part, lds);
}

// ---------------------------------------------------------------------------
// small dense solve in one lane (Gaussian elimination, partial pivoting)
// M is n x (n+1) augmented, row stride ld, in LDS or private memory.
// ---------------------------------------------------------------------------
__device__ inline bool dense_solve(double *M, int n, int ld, double *x) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = fabs(M[c * ld + c]);
    for (int r = c + 1; r < n; ++r) {
      const double v = fabs(M[r * ld + c]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (p != c) {
      for (int j = 0; j <= n; ++j) {
        const double tmp = M[c * ld + j];
        M[c * ld + j] = M[p * ld + j];
        M[p * ld + j] = tmp;
      }
    }
    const double piv = M[c * ld + c];
    if (piv == 0.0) return false;
    for (int r = c + 1; r < n; ++r) {
      const double f = M[r * ld + c] / piv;
      if (f != 0.0) {
        for (int j = c; j <= n; ++j) M[r * ld + j] -= f * M[c * ld + j];
      }
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = 0.0;
    for (int j = r + 1; j < n; ++j) s += M[r * ld + j] * x[j];
    x[r] = (M[r * ld + n] - s) / M[r * ld + r];
  }
  return true;
}

// Primal active-set QP: min 1/2 x'Px + q'x, x >= 0, sum x = pop (same steps
// as the oracle's qp_simplex_active_set).  Returns false if not converged.
__device__ inline bool qp_simplex(const double *P, const double *q, double pop, int K, double *x, double *M) {
  bool fixed[HP_KMAX];
  int freeidx[HP_KMAX];
  double xs[HP_KMAX + 1];
  for (int k = 0; k < K; ++k) {
    x[k] = pop / K;
    fixed[k] = false;
  }
  for (int it = 0; it < 200; ++it) {
    int mfree = 0;
    for (int k = 0; k < K; ++k)
      if (!fixed[k]) freeidx[mfree++] = k;
    const int n = mfree + 1, ld = HP_KMAX + 2;
    for (int i = 0; i < mfree; ++i) {
      for (int j = 0; j < mfree; ++j) M[i * ld + j] = P[freeidx[i] * K + freeidx[j]];
      M[i * ld + mfree] = -1.0;
      M[i * ld + n] = -q[freeidx[i]];
    }
    for (int j = 0; j < mfree; ++j) M[mfree * ld + j] = 1.0;
    M[mfree * ld + mfree] = 0.0;
    M[mfree * ld + n] = pop;
    if (!dense_solve(M, n, ld, xs)) return false;
    bool feasible = true;
    for (int i = 0; i < mfree; ++i)
      if (!(xs[i] >= 0.0)) feasible = false;
    if (feasible) {
      const double nu = xs[mfree];
      for (int k = 0; k < K; ++k) x[k] = 0.0;
      for (int i = 0; i < mfree; ++i) x[freeidx[i]] = xs[i];
      int jmin = -1;
      double mumin = 0.0;
      bool anyfixed = false;
      for (int k = 0; k < K; ++k) {
        double mu = 0.0;
        if (fixed[k]) {
          anyfixed = true;
          double g = q[k];
          for (int j = 0; j < K; ++j) g += P[k * K + j] * x[j];
          mu = g - nu;
        }
        if (jmin < 0 || mu < mumin) {
          mumin = mu;
          jmin = k;
        }
      }
      if (!anyfixed || mumin >= 0.0) return true;
      fixed[jmin] = false;
    } else {
      double alpha = 1.0;
      int block = -1;
      for (int i = 0; i < mfree; ++i) {
        const int k = freeidx[i];
        if (xs[i] < 0.0) {
          const double al = x[k] / (x[k] - xs[i]);
          if (al < alpha) {
            alpha = al;
            block = k;
          }
        }
      }
      double xn[HP_KMAX];
      for (int k = 0; k < K; ++k) xn[k] = 0.0;
      for (int i = 0; i < mfree; ++i) {
        const int k = freeidx[i];
        xn[k] = x[k] + alpha * (xs[i] - x[k]);
      }
      for (int k = 0; k < K; ++k) x[k] = xn[k];
      if (block >= 0) {
        x[block] = 0.0;
        fixed[block] = true;
      }
    }
  }
  return false;
}

// Basis value n alpha^(3/n)/(4 pi Gamma(3/n)) exp(-alpha r^n) for a radial point.
__device__ __forceinline__ double basis_radial(double norm, double alpha, double order, double r) {
  const double arg = (order == 1.0) ? r : ((order == 2.0) ? r * r : pow(r, order));
  return norm * exp(-alpha * arg);
}

// ---------------------------------------------------------------------------
// K4: GISA QP update, one wavefront per atom (gisa.py:281-318, 348-421)
// ---------------------------------------------------------------------------
template <int PPL>
__global__ __launch_bounds__(256) void k_gisa_qp(const int *__restrict__ done, DevMol m, const double *__restrict__ rhobar, double *__restrict__ par,
                                                 const double *__restrict__ old_par, const int *__restrict__ fn_off,
                                                 const double *__restrict__ fn_norm, const double *__restrict__ fn_alpha,
                                                 const double *__restrict__ fn_order, const double *__restrict__ qpP,
                                                 const int64_t *__restrict__ qp_off, double pop_cutoff, double neg_cutoff,
                                                 int *__restrict__ inner_niter, int *__restrict__ status,
                                                 double *__restrict__ charges, double *__restrict__ change_part) {
  __shared__ double Mlds[4][(HP_KMAX + 1) * (HP_KMAX + 2)];
  if (done && *done) return;
  __builtin_amdgcn_s_setprio(3);  // latency-bound: issue first (see k_mbis_inner)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ai = blockIdx.x * (blockDim.x >> 6) + wv;
  if (ai >= m.natom) return;
  int a = ai;
  if (m.work) {
    const int2 wk = m.work[ai];
    if (m.slots[wk.y].run == HP_RUN_IDLE) return;
    a = wk.x;
  }
  const int s0 = m.atom_shell_off[a], R = m.atom_shell_off[a + 1] - s0;
  const int f0 = fn_off[a], K = fn_off[a + 1] - f0;
  RadialLane<PPL> d;
  load_radial<PPL>(m, rhobar, s0, R, lane, d);
  double red[HP_KMAX + 1];
  for (int k = 0; k <= HP_KMAX; ++k) red[k] = 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    for (int k = 0; k < K; ++k) {
      const double B = basis_radial(fn_norm[f0 + k], fn_alpha[f0 + k], fn_order[f0 + k], d.r[j]);
      red[k] += d.w[j] * B * d.rho[j];
    }
    red[HP_KMAX] += d.w[j] * d.rho[j];
  }
  wave_sum_n<HP_KMAX + 1>(red);
  const double pop = red[HP_KMAX];
  double x[HP_KMAX];
  for (int k = 0; k < HP_KMAX; ++k) x[k] = 0.0;
  int st = 0;
  if (lane == 0) {
    double q[HP_KMAX];
    for (int k = 0; k < K; ++k) q[k] = -2.0 * red[k];
    const bool ok = qp_simplex(qpP + qp_off[a], q, pop, K, x, Mlds[wv]);
    if (!ok) st |= HP_ST_QPFAIL;
    double sx = 0.0;
    for (int k = 0; k < K; ++k) {
      if (x[k] < neg_cutoff) st |= HP_ST_NEGPARS;
      sx += x[k];
    }
    if (fabs(sx - pop) > pop_cutoff) st |= HP_ST_POP;
    for (int k = 0; k < K; ++k) par[f0 + k] = x[k];
  }
  for (int k = 0; k < HP_KMAX; ++k) x[k] = __shfl(k < K ? x[k] : 0.0, 0, 64);
  // outer change on the radial grid
  double msd = 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    double y1 = 0.0, y2 = 0.0;
    for (int k = 0; k < K; ++k) {
      const double B = basis_radial(fn_norm[f0 + k], fn_alpha[f0 + k], fn_order[f0 + k], d.r[j]);
      y1 += B * x[k];
      y2 += B * old_par[f0 + k];
    }
    const double e = y1 - y2;
    msd += d.wr[j] * (4.0 * HP_PI * (d.r[j] * d.r[j])) * e * e;
  }
  msd = wave_sum(msd);
  if (lane == 0) {
    inner_niter[a] = 1;
    status[a] = st;
    charges[a] = m.pseudo[a] - pop;
    change_part[a] = msd;
  }
}

// ---------------------------------------------------------------------------
// K5: LISA solvers on the radial grid, one wavefront per atom
// (alisa.py:193-291 solver_sc, 753-938 _solver_general_newton)
// ---------------------------------------------------------------------------

// For all points of the wave: 1 if any value v[j] (valid points only) fails
// v[j] > cut, or, with mono, v[i] - v[i+1] > cut fails for a consecutive pair.
template <int PPL>
__device__ __forceinline__ bool wave_any_bad(const double *v, int R, int lane, double cut, bool mono) {
  double bad = 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int i = lane + 64 * j;
    if (i < R && !(v[j] > cut)) bad = 1.0;
  }
  if (mono) {
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const int i = lane + 64 * j;
      double nxt = __shfl_down(v[j], 1, 64);
      if (j + 1 < PPL) {
        const double first_next = __shfl(v[j + 1 < PPL ? j + 1 : j], 0, 64);
        if (lane == 63) nxt = first_next;
      }
      if (i + 1 < R && !(v[j] - nxt > cut)) bad = 1.0;
    }
  }
  return wave_sum(bad) > 0.0;
}

template <int PPL, int KM>
__global__ __launch_bounds__(256) void k_lisa_inner(const int *__restrict__ done, DevMol m, const double *__restrict__ rhobar, double *__restrict__ par,
                                                    const double *__restrict__ old_par, const int *__restrict__ fn_off,
                                                    const double *__restrict__ fn_norm, const double *__restrict__ fn_alpha,
                                                    const double *__restrict__ fn_order, int solver, double thr, int maxit,
                                                    double dc, double neg_cutoff, double pop_cutoff, double tau,
                                                    int nlin, int check_mono, int *__restrict__ inner_niter,
                                                    int *__restrict__ status, double *__restrict__ charges,
                                                    double *__restrict__ change_part) {
  __shared__ double Mlds[4][(HP_KMAX + 1) * (HP_KMAX + 2)];
  if (done && *done) return;
  __builtin_amdgcn_s_setprio(3);  // latency-bound: issue first (see k_mbis_inner)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ai = blockIdx.x * (blockDim.x >> 6) + wv;
  if (ai >= m.natom) return;
  int a = ai;
  if (m.work) {
    const int2 wk = m.work[ai];
    if (m.slots[wk.y].run == HP_RUN_IDLE) return;
    a = wk.x;
  }
  const int s0 = m.atom_shell_off[a], R = m.atom_shell_off[a + 1] - s0;
  const int f0 = fn_off[a], K = fn_off[a + 1] - f0;
  RadialLane<PPL> d;
  load_radial<PPL>(m, rhobar, s0, R, lane, d);
  double B[KM][PPL];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
#pragma unroll
    for (int j = 0; j < PPL; ++j)
      B[k][j] = (k < K && lane + 64 * j < R) ? basis_radial(fn_norm[f0 + k], fn_alpha[f0 + k], fn_order[f0 + k], d.r[j])
                                              : 0.0;
  }
  double pop = 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) pop += d.w[j] * d.rho[j];
  pop = wave_sum(pop);
  double c[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) c[k] = k < K ? par[f0 + k] : 0.0;
  double oldpro[PPL];
#pragma unroll
  for (int j = 0; j < PPL; ++j) oldpro[j] = 0.0;
  bool have_old = false, conv = false;
  int st = 0, it = 0;
  double change = 1e100;
  for (it = 0; it < maxit; ++it) {
    double pro[PPL], ratio[PPL];
    bool sick[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < KM; ++k)
        if (k < K) s += B[k][j] * c[k];
      pro[j] = s;
      sick[j] = (d.rho[j] < dc) || (s < dc);
      ratio[j] = sick[j] ? 0.0 : d.rho[j] / s;
    }
    if (solver == 2) {
      // sc: c_k <- sum_p w_p (B_kp c_k) ratio_p
      double red[KM + 1];
#pragma unroll
      for (int k = 0; k <= KM; ++k) red[k] = 0.0;
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
#pragma unroll
        for (int k = 0; k < KM; ++k) red[k] += d.w[j] * ((B[k][j] * c[k]) * ratio[j]);
        const double e = oldpro[j] - pro[j];
        red[KM] += d.w[j] * e * e;
      }
      wave_sum_n<KM + 1>(red);
#pragma unroll
      for (int k = 0; k < KM; ++k)
        if (k < K) c[k] = red[k];
      change = have_old ? sqrt(red[KM]) : 1e100;
      if (change < thr) {
        conv = true;
        break;
      }
    } else {
      if (have_old) {
        double e2 = 0.0;
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          const double e = oldpro[j] - pro[j];
          e2 += d.w[j] * e * e;
        }
        change = sqrt(wave_sum(e2));
      }
      if (change < thr) {
        conv = true;
        // check_pro_atom_parameters_neg_pars (alisa.py:895-904)
        if (wave_any_bad<PPL>(pro, R, lane, neg_cutoff, false)) {
          double mn = 1e300;
#pragma unroll
          for (int j = 0; j < PPL; ++j)
            if (lane + 64 * j < R) mn = fmin(mn, pro[j]);
          if (wave_min(mn) < neg_cutoff) st |= HP_ST_NEGDENS;
        }
        if (check_mono) {
          // monotonic decay: pro[i] - pro[i+1] >= neg_cutoff (utils.py:546)
          double bad = 0.0;
#pragma unroll
          for (int j = 0; j < PPL; ++j) {
            const int i = lane + 64 * j;
            double nxt = __shfl_down(pro[j], 1, 64);
            if (j + 1 < PPL) {
              const double first_next = __shfl(pro[j + 1 < PPL ? j + 1 : j], 0, 64);
              if (lane == 63) nxt = first_next;
            }
            if (i + 1 < R && (pro[j] - nxt < neg_cutoff)) bad = 1.0;
          }
          if (wave_sum(bad) > 0.0) st |= HP_ST_NOTMONO;
        }
        break;
      }
      // gradient and Hessian (alisa.py:907-924)
      constexpr int NH = KM * (KM + 1) / 2;
      double red[KM + NH];
#pragma unroll
      for (int q = 0; q < KM + NH; ++q) red[q] = 0.0;
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const double integ = B[k][j] * ratio[j];
          red[k] += integ * d.w[j];
          const double gk = sick[j] ? 0.0 : integ / pro[j];
#pragma unroll
          for (int l = 0; l <= k; ++l) red[KM + k * (k + 1) / 2 + l] += gk * B[l][j] * d.w[j];
        }
      }
      wave_sum_n<KM + NH>(red);
      bool finite = true;
#pragma unroll
      for (int q = 0; q < KM + NH; ++q) finite = finite && isfinite(red[q]);
      if (!finite) {
        // scipy.linalg.solve(check_finite=True) raises ValueError
        st |= HP_ST_NONFINITE;
        break;
      }
      double delta[KM];
      int solved = 1;
      if (lane == 0) {
        double *M = Mlds[wv];
        const int ld = HP_KMAX + 2;
#pragma unroll
        for (int k = 0; k < KM; ++k) {
#pragma unroll
          for (int l = 0; l <= k; ++l) {
            if (k < K) {
              M[k * ld + l] = red[KM + k * (k + 1) / 2 + l];
              M[l * ld + k] = red[KM + k * (k + 1) / 2 + l];
            }
          }
          if (k < K) M[k * ld + K] = -1.0 - (-red[k]);
        }
        double x[HP_KMAX];
        for (int k = 0; k < HP_KMAX; ++k) x[k] = 0.0;
        solved = dense_solve(M, K, ld, x) ? 1 : 0;
#pragma unroll
        for (int k = 0; k < KM; ++k) delta[k] = k < K ? x[k] : 0.0;
      }
      if (!__shfl(solved, 0, 64)) {
        // exactly singular Hessian: scipy.linalg.solve raises LinAlgError
        st |= HP_ST_SINGULAR;
        break;
      }
#pragma unroll
      for (int k = 0; k < KM; ++k) delta[k] = __shfl(lane == 0 ? delta[k] : 0.0, 0, 64);
      if (solver == 3) {
#pragma unroll
        for (int k = 0; k < KM; ++k) c[k] = c[k] + delta[k];
      } else {
        // modified: first x in linspace(tau, 0, nlin) keeping pro > neg_cutoff
        const double step = (0.0 - tau) / (double)(nlin - 1);
        bool found = false;
        for (int t = 0; t < nlin && !found; ++t) {
          const double xs = (t == nlin - 1) ? 0.0 : (double)t * step + tau;
          double cn[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k) cn[k] = c[k] + xs * delta[k];
          double np_[PPL];
#pragma unroll
          for (int j = 0; j < PPL; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < KM; ++k)
              if (k < K) s += B[k][j] * cn[k];
            np_[j] = s;
          }
          if (!wave_any_bad<PPL>(np_, R, lane, neg_cutoff, check_mono != 0)) {
            found = true;
#pragma unroll
            for (int k = 0; k < KM; ++k) c[k] = cn[k];
          }
        }
        if (!found) {
          st |= HP_ST_LINESEARCH;
          break;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < PPL; ++j) oldpro[j] = pro[j];
    have_old = true;
  }
  if (!conv) {
    if (solver == 2)
      st |= HP_ST_NOCONV;
    else if (!(st & (HP_ST_LINESEARCH | HP_ST_SINGULAR | HP_ST_NONFINITE)))
      st |= HP_ST_NEWTONMAX;
  } else {
    double sx = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) sx += c[k];
    if (fabs(sx - pop) > pop_cutoff) st |= HP_ST_POP;
  }
  double msd = 0.0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    double y1 = 0.0, y2 = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      if (k < K) {
        y1 += B[k][j] * c[k];
        y2 += B[k][j] * old_par[f0 + k];
      }
    }
    const double e = y1 - y2;
    msd += d.wr[j] * (4.0 * HP_PI * (d.r[j] * d.r[j])) * e * e;
  }
  msd = wave_sum(msd);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) par[f0 + k] = c[k];
    inner_niter[a] = conv ? it + 1 : it;
    status[a] = st;
    charges[a] = m.pseudo[a] - pop;
    change_part[a] = msd;
  }
}

// ---------------------------------------------------------------------------
// K3: per-iteration finalisation (single workgroup): entropy and change sums,
// histories, and the stop decision of iterstock.py:187 on the device:
// stop when (iteration >= min_stop and change < thr) or iteration >= max_iter
// or a per-atom solver raised.  state[0] = done, state[1] = iterations so far.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(256) void k_finalize(int *__restrict__ state, double thr, int min_stop, int max_iter,
                                                  int nshell, const double *__restrict__ ent_part, int natom,
                                                  const double *__restrict__ change_part, const int *__restrict__ status,
                                                  int *__restrict__ status_all, double *__restrict__ par, int npar,
                                                  const double *__restrict__ charges, double *__restrict__ hpar,
                                                  double *__restrict__ hch, double *__restrict__ hent,
                                                  double *__restrict__ hchg, double *__restrict__ out,
                                                  const double *__restrict__ old_par, int rollback) {
  __shared__ double lds[32];
  __shared__ int s_rollback;
  if (state[0]) return;
  const int iter = state[1];
  double e = 0.0, dummy = 0.0;
  for (int s = threadIdx.x; s < nshell; s += blockDim.x) e += ent_part[s];
  block_sum2(e, dummy, lds);
  for (int i = threadIdx.x; i < npar; i += blockDim.x) hpar[(int64_t)iter * npar + i] = par[i];
  for (int i = threadIdx.x; i < natom; i += blockDim.x) {
    hch[(int64_t)iter * natom + i] = charges[i];
    status_all[i] |= status[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double msd = 0.0;
    int err = 0, any = 0;
    for (int a = 0; a < natom; ++a) {
      msd += change_part[a];
      err |= status[a] & HP_ST_FATAL;
      any |= status[a];
    }
    const double ch = sqrt(msd);
    hent[iter] = e;
    hchg[iter] = ch;
    out[0] = ch;
    out[1] = e;
    out[2] = (double)err;
    out[3] = (double)any;
    const int n = iter + 1;
    state[1] = n;
    s_rollback = 0;
    if (err != 0 || n >= max_iter || (n >= min_stop && ch < thr)) {
      state[0] = 1;
      // the iteration that ends a stage with rollback is discarded: the
      // next stage redoes it from the parameters it started from
      if (rollback && err == 0) {
        state[1] = iter;
        s_rollback = 1;
      }
    }
  }
  __syncthreads();
  if (s_rollback)
    for (int i = threadIdx.x; i < npar; i += blockDim.x) par[i] = old_par[i];
}

// ---------------------------------------------------------------------------
// moments up to lmax = 3: one workgroup per radial shell, 40 partial sums
// ---------------------------------------------------------------------------
#define HP_NMOM 40
#define HP_MOM_CHUNK 4096  // molecular-grid moments: points per workgroup

// Adds wf times the 40 moment integrands of the offset (x, y, z) from the
// centre: 20 cartesian powers (HORTON order), 16 real regular solid harmonics
// (Racah normalisation; c_l0, c_l1, s_l1, ...), radial powers r^0..r^3.
__device__ __forceinline__ void add_moment_terms(double wf, double x, double y, double z, double *acc) {
  const double xx = x * x, yy = y * y, zz = z * z, r2 = xx + yy + zz;
  const double c[20] = {1.0,        x,          y,          z,          xx,         x * y,      x * z,
                        yy,         y * z,      zz,         xx * x,     xx * y,     xx * z,     x * yy,
                        x * y * z,  x * zz,     yy * y,     yy * z,     y * zz,     zz * z};
  for (int i = 0; i < 20; ++i) acc[i] += wf * c[i];
  const double s3 = 1.7320508075688772, s15 = 3.872983346207417;
  const double s38 = 0.6123724356957945, s58 = 0.7905694150420949;
  const double pu[16] = {1.0,
                         z,
                         x,
                         y,
                         0.5 * (3.0 * zz - r2),
                         s3 * x * z,
                         s3 * y * z,
                         0.5 * s3 * (xx - yy),
                         s3 * x * y,
                         0.5 * z * (5.0 * zz - 3.0 * r2),
                         s38 * x * (5.0 * zz - r2),
                         s38 * y * (5.0 * zz - r2),
                         0.5 * s15 * z * (xx - yy),
                         s15 * x * y * z,
                         s58 * x * (xx - 3.0 * yy),
                         s58 * y * (3.0 * xx - yy)};
  for (int i = 0; i < 16; ++i) acc[20 + i] += wf * pu[i];
  const double r = sqrt(r2);
  acc[36] += wf;
  acc[37] += wf * r;
  acc[38] += wf * r2;
  acc[39] += wf * r2 * r;
}

// Sums the 40 accumulators of a 256-thread workgroup (waves in order) into out.
__device__ __forceinline__ void block_moments_store(double *acc, double *lds, double *__restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i = 0; i < HP_NMOM; ++i) {
    double v = acc[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) lds[i * 4 + wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < HP_NMOM) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += lds[threadIdx.x * 4 + w];
    out[threadIdx.x] = v;
  }
}

// moments (lmax <= 3) on atom grids: one workgroup per radial shell, the
// shell atom's AIM density with the atom-grid weights
template <int FK>
__global__ __launch_bounds__(256) void k_moments(DevMol m, DevFn fn, double *__restrict__ part) {
  __shared__ double lds[HP_NMOM * 4];
  const int s = blockIdx.x;
  const int a = m.shell_atom[s];
  const int64_t p0 = m.shell_off[s], p1 = m.shell_off[s + 1];
  double acc[HP_NMOM];
  for (int i = 0; i < HP_NMOM; ++i) acc[i] = 0.0;
  DevFn f2 = fn;
  f2.nb_list = nullptr;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    double promol, own;
    promol_at<double, FK>(m, f2, a, m.natom, nullptr, m.px[p], m.py[p], m.pz[p], promol, own);
    const double w = aim_weight(own, promol);
    const double wf = m.wat[p] * (m.rho[p] * w);
    add_moment_terms(wf, m.px[p] - m.cx[a], m.py[p] - m.cy[a], m.pz[p] - m.cz[a], acc);
  }
  block_moments_store(acc, lds, part + (int64_t)s * HP_NMOM);
}

// moments on the molecular grid (grid_type 3): workgroup (chunk, atom) over
// HP_MOM_CHUNK points, the atom's AIM density from the stored promolecule
// with the molecular weights
template <int FK>
__global__ __launch_bounds__(256) void k_moments_mol(DevMol m, DevFn fn, const double *__restrict__ promol,
                                                     double *__restrict__ part) {
  __shared__ double lds[HP_NMOM * 4];
  const int a = blockIdx.y;
  const int64_t p0 = (int64_t)blockIdx.x * HP_MOM_CHUNK;
  const int64_t p1 = p0 + HP_MOM_CHUNK < m.npoint ? p0 + HP_MOM_CHUNK : m.npoint;
  double acc[HP_NMOM];
  for (int i = 0; i < HP_NMOM; ++i) acc[i] = 0.0;
  const bool cull = fn.rc > 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const double x = m.px[p] - m.cx[a], y = m.py[p] - m.cy[a], z = m.pz[p] - m.cz[a];
    const double r2 = x * x + y * y + z * z;
    const double r = sqrt(r2);
    const double own = (!cull || r <= fn.rc) ? proatom_value<double, FK>(fn, a, r, r2) : 0.0;
    const double wf = m.wmol[p] * (m.rho[p] * aim_weight(own, promol[p]));
    add_moment_terms(wf, x, y, z, acc);
  }
  block_moments_store(acc, lds, part + ((int64_t)a * gridDim.x + blockIdx.x) * HP_NMOM);
}

static __global__ void k_moments_reduce_mol(int nchunk, const double *__restrict__ part, double *__restrict__ out) {
  const int a = blockIdx.x;
  const int i = threadIdx.x;
  if (i >= HP_NMOM) return;
  double v = 0.0;
  for (int c = 0; c < nchunk; ++c) v += part[((int64_t)a * nchunk + c) * HP_NMOM + i];
  out[(int64_t)a * HP_NMOM + i] = v;
}

// Final layout of the moments from the per-atom sums (rows of `stride`):
// cart (natom x ncart), pure (natom x npure), radial (natom x nrad) one after
// the other, taken from the row offsets src_*; the electron sign applied and
// pseudo_numbers added to the monopoles (core/base.py:374-402).
static __global__ void k_moments_layout(int natom, int stride, int src_cart, int ncart, int src_pure, int npure,
                                        int src_rad, int nrad, const double *__restrict__ mom,
                                        const double *__restrict__ pseudo, double *__restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natom) return;
  const double *v = mom + (int64_t)a * stride;
  double *cart = out + (int64_t)a * ncart, *pure = out + (int64_t)natom * ncart + (int64_t)a * npure,
         *rad = out + (int64_t)natom * (ncart + npure) + (int64_t)a * nrad;
  for (int i = 0; i < ncart; ++i) cart[i] = -v[src_cart + i];
  cart[0] += pseudo[a];
  for (int i = 0; i < npure; ++i) pure[i] = -v[src_pure + i];
  pure[0] += pseudo[a];
  for (int i = 0; i < nrad; ++i) rad[i] = v[src_rad + i];
}

// Sum of the per-shell partials (one workgroup of 256 threads): the entropy
// of a weights pass, in the order k_finalize sums it.
static __global__ __launch_bounds__(256) void k_sum_shells(int nshell, const double *__restrict__ part,
                                                           double *__restrict__ out) {
  __shared__ double lds[32];
  double e = 0.0, dummy = 0.0;
  for (int s = threadIdx.x; s < nshell; s += blockDim.x) e += part[s];
  block_sum2(e, dummy, lds);
  if (threadIdx.x == 0) *out = e;
}

static __global__ void k_moments_reduce(DevMol m, const double *__restrict__ part, double *__restrict__ out) {
  const int a = blockIdx.x;
  const int i = threadIdx.x;
  if (i >= HP_NMOM) return;
  double v = 0.0;
  for (int s = m.atom_shell_off[a]; s < m.atom_shell_off[a + 1]; ++s) v += part[(int64_t)s * HP_NMOM + i];
  out[(int64_t)a * HP_NMOM + i] = v;
}

// ---------------------------------------------------------------------------
// moments of any order: one lane per quantity
// ---------------------------------------------------------------------------
// The quantities of order lmax (decode_moment_quantity, moment_quantity): a
// workgroup covers HP_MOM_GROUP of them (blockIdx.z), two per lane; its
// points are staged in LDS 256 at a time and each wave takes every fourth
// staged point.
#define HP_MOM_GROUP 128

// Moments of order lmax (any): MOL = false, one workgroup per radial shell
// (the shell atom's AIM density with the atom-grid weights, as k_moments);
// MOL = true, workgroup (point chunk, atom) with the molecular weights and the
// stored promolecule (as k_moments_mol).  Row of part: the shell, or
// atom * nchunk + chunk; nq quantities per row.
template <int FK, bool MOL>
__global__ __launch_bounds__(256) void k_moments_any(DevMol m, DevFn fn, const double *__restrict__ promol, int lmax,
                                                     int nq, double *__restrict__ part) {
  __shared__ double s_w[256], s_x[256], s_y[256], s_z[256];
  __shared__ double red[4][HP_MOM_GROUP];
  int a;
  int64_t p0, p1, row;
  if (MOL) {
    a = blockIdx.y;
    p0 = (int64_t)blockIdx.x * HP_MOM_CHUNK;
    p1 = p0 + HP_MOM_CHUNK < m.npoint ? p0 + HP_MOM_CHUNK : m.npoint;
    row = (int64_t)a * gridDim.x + blockIdx.x;
  } else {
    a = m.shell_atom[blockIdx.x];
    p0 = m.shell_off[blockIdx.x];
    p1 = m.shell_off[blockIdx.x + 1];
    row = blockIdx.x;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int qa = (int)blockIdx.z * HP_MOM_GROUP + lane, qb = qa + 64;
  const MomQuantity da = decode_moment_quantity(qa < nq ? qa : -1, lmax);
  const MomQuantity db = decode_moment_quantity(qb < nq ? qb : -1, lmax);
  DevFn f2 = fn;
  f2.nb_list = nullptr;
  const bool cull = fn.rc > 0.0;
  double acc_a = 0.0, acc_b = 0.0;
  for (int64_t base = p0; base < p1; base += 256) {
    const int64_t p = base + threadIdx.x;
    double wf = 0.0, x = 0.0, y = 0.0, z = 0.0;
    if (p < p1) {
      x = m.px[p] - m.cx[a];
      y = m.py[p] - m.cy[a];
      z = m.pz[p] - m.cz[a];
      if (MOL) {
        const double r2 = x * x + y * y + z * z;
        const double r = sqrt(r2);
        const double own = (!cull || r <= fn.rc) ? proatom_value<double, FK>(fn, a, r, r2) : 0.0;
        wf = m.wmol[p] * (m.rho[p] * aim_weight(own, promol[p]));
      } else {
        double pm, own;
        promol_at<double, FK>(m, f2, a, m.natom, nullptr, m.px[p], m.py[p], m.pz[p], pm, own);
        wf = m.wat[p] * (m.rho[p] * aim_weight(own, pm));
      }
    }
    s_w[threadIdx.x] = wf;
    s_x[threadIdx.x] = x;
    s_y[threadIdx.x] = y;
    s_z[threadIdx.x] = z;
    __syncthreads();
    const int cnt = p1 - base < 256 ? (int)(p1 - base) : 256;
    for (int j = wv; j < cnt; j += 4) {
      const double w = s_w[j], xj = s_x[j], yj = s_y[j], zj = s_z[j];
      if (da.kind >= 0) acc_a += w * moment_quantity(da, xj, yj, zj);
      if (db.kind >= 0) acc_b += w * moment_quantity(db, xj, yj, zj);
    }
    __syncthreads();
  }
  red[wv][lane] = acc_a;
  red[wv][lane + 64] = acc_b;
  __syncthreads();
  if (threadIdx.x < HP_MOM_GROUP) {
    const int q = (int)blockIdx.z * HP_MOM_GROUP + threadIdx.x;
    if (q < nq) part[row * nq + q] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
  }
}

// Per-atom sums of the partial rows (nq per row): over the atom's shells
// (atom grids) or its nchunk point chunks (molecular grid), in row order.
static __global__ void k_moments_reduce_any(DevMol m, int molgrid, int nchunk, int nq, const double *__restrict__ part,
                                            double *__restrict__ out) {
  const int a = blockIdx.x;
  const int64_t r0 = molgrid ? (int64_t)a * nchunk : m.atom_shell_off[a];
  const int64_t r1 = molgrid ? r0 + nchunk : m.atom_shell_off[a + 1];
  for (int i = threadIdx.x; i < nq; i += blockDim.x) {
    double v = 0.0;
    for (int64_t r = r0; r < r1; ++r) v += part[r * nq + i];
    out[(int64_t)a * nq + i] = v;
  }
}

// ---------------------------------------------------------------------------
// grid utilities
// ---------------------------------------------------------------------------

// Becke fuzzy-cell weight of each point for the atom whose grid holds it
// (A. D. Becke, J. Chem. Phys. 88 (1988) 2547; size-adjusted, order-3 step).
// adj[i*A+j] = a_ij, invr[i*A+j] = 1/|R_i - R_j|, ord[o*A + k] = the atoms
// by distance from atom o (o itself first).
//
// Per point the cell products P_i = prod_{j != i} s(nu_ij) run over the atoms
// in the order of the point's owner: the atoms nearest the point come first,
// so the product of an atom far from the point falls fast.  Every factor is
// at most 1, so once P_i is at most 2^-64 of the sum of the products before
// it, P_i stays there: the product stops and the atom adds nothing.  Each
// omitted term is below 2^-64 of the final sum (natom of them move the
// weight by less than natom x 2^-64 relative).  The owner comes first, with
// an empty sum, and its product stops only at 0, as every product did before.
// The distances of the first HP_BECKE_NC atoms of the order are formed once
// per point.
#define HP_BECKE_NC 8

__device__ __forceinline__ double becke_factor(double di, double dj, double invr_ij, double adj_ij, int order) {
  const double mu = (di - dj) * invr_ij;
  double f = mu + adj_ij * (1.0 - mu * mu);
  for (int k = 0; k < order; ++k) f = 1.5 * f - 0.5 * (f * f * f);
  return 0.5 * (1.0 - f);
}

// One point's weight.  UNI: the owner is the same on every lane of the
// wavefront (passed as a wave-uniform value), so the order, the coordinates
// and the table rows are read with scalar loads.
template <bool UNI>
__device__ __forceinline__ double becke_point(int natom, const double *__restrict__ cx, const double *__restrict__ cy,
                                              const double *__restrict__ cz, const double *__restrict__ adj,
                                              const double *__restrict__ invr, const int *__restrict__ ord, int lo,
                                              double x, double y, double z, int order) {
  constexpr int NC = HP_BECKE_NC;
  const int *o = ord + (int64_t)lo * natom;
  double dc[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    dc[k] = 0.0;
    if (k < natom) {
      const int j = UNI ? __builtin_amdgcn_readfirstlane(o[k]) : o[k];
      const double dx = x - cx[j], dy = y - cy[j], dz = z - cz[j];
      dc[k] = sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  double sum = 0.0, own = 0.0;
  for (int ki = 0; ki < natom; ++ki) {
    const int i = UNI ? __builtin_amdgcn_readfirstlane(o[ki]) : o[ki];
    double di;
    if (ki < NC) {
      di = dc[0];
#pragma unroll
      for (int k = 1; k < NC; ++k) di = ki == k ? dc[k] : di;
    } else {
      const double dx = x - cx[i], dy = y - cy[i], dz = z - cz[i];
      di = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const double thr = 0x1p-64 * sum;
    const double *ai = adj + (int64_t)i * natom, *ri = invr + (int64_t)i * natom;
    double P = 1.0;
    bool stop = false;
#pragma unroll
    for (int kj = 0; kj < NC; ++kj) {
      if (kj < natom && !stop) {
        const int j = UNI ? __builtin_amdgcn_readfirstlane(o[kj]) : o[kj];
        if (j != i) {
          P *= becke_factor(di, dc[kj], ri[j], ai[j], order);
          stop = P <= thr;
        }
      }
    }
    if (!stop) {
      for (int kj = NC; kj < natom; ++kj) {
        const int j = UNI ? __builtin_amdgcn_readfirstlane(o[kj]) : o[kj];
        if (j == i) continue;
        const double dx = x - cx[j], dy = y - cy[j], dz = z - cz[j];
        const double dj = sqrt(dx * dx + dy * dy + dz * dz);
        P *= becke_factor(di, dj, ri[j], ai[j], order);
        if (P <= thr) {
          stop = true;
          break;
        }
      }
    }
    if (ki == 0) {
      own = P;
      sum = P;
    } else if (!stop) {
      sum += P;
    }
  }
  return own / sum;
}

static __global__ void k_becke(int natom, const double *__restrict__ cx, const double *__restrict__ cy,
                               const double *__restrict__ cz, const double *__restrict__ adj,
                               const double *__restrict__ invr, const int *__restrict__ ord, int64_t npoint,
                               const double *__restrict__ px, const double *__restrict__ py,
                               const double *__restrict__ pz, const int64_t *__restrict__ atom_off, int order,
                               double *__restrict__ out) {
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = p0 < npoint;
  const int64_t p = ok ? p0 : npoint - 1;  // idle lanes follow the last point, so they stay on its owner
  int lo = 0, hi = natom;  // owner: atom_off[lo] <= p < atom_off[lo+1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (atom_off[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  const double x = px[p], y = py[p], z = pz[p];
  double w = 1.0;
  if (natom > 1) {
    const int lo_u = __builtin_amdgcn_readfirstlane(lo);
    if (__all(lo == lo_u))
      w = becke_point<true>(natom, cx, cy, cz, adj, invr, ord, lo_u, x, y, z, order);
    else
      w = becke_point<false>(natom, cx, cy, cz, adj, invr, ord, lo, x, y, z, order);
  }
  if (ok) out[p] = w;
}

// sum_a sum_k coef_ak exp(-alpha_ak |p - R_a|) at every point, one
// workgroup per 256 consecutive points.  With cull (every coefficient and
// exponent non-negative, so every term decreases with distance) the
// workgroup bounds each atom over its points' bounding box first: atom a adds
// at least lo_a = max_k c_ak e^(-alpha_ak dmax_a) and at most
// hi_a = nf_a max_k c_ak e^(-alpha_ak dmin_a) at every point of the box
// (dmin, dmax the nearest and farthest distances from R_a to the box).  An
// atom with hi_a < 2^-64 max_b lo_b adds less than 2^-64 of the density at
// every point and is skipped.  The test runs in the log2 domain in FP32 with
// two units of margin and 1e-6 relative margins on the distances; the kept
// atoms are summed in atom order with the FP64 terms of the plain sum.
static __global__ __launch_bounds__(256) void k_slater_sum(int natom, const double *__restrict__ cx,
                                                           const double *__restrict__ cy, const double *__restrict__ cz,
                                                           const int *__restrict__ fn_off,
                                                           const double *__restrict__ coef,
                                                           const double *__restrict__ alpha, int64_t npoint,
                                                           const double *__restrict__ px, const double *__restrict__ py,
                                                           const double *__restrict__ pz, int cull,
                                                           double *__restrict__ out) {
  __shared__ float s_red[6][4];
  __shared__ float s_lo[4];
  __shared__ int s_cnt[4];
  __shared__ int s_keep[256];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = p < npoint;
  const int64_t pc = ok ? p : npoint - 1;
  const double x = px[pc], y = py[pc], z = pz[pc];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double acc = 0.0;
  if (!cull) {
    for (int a = 0; a < natom; ++a) {
      const double dx = x - cx[a], dy = y - cy[a], dz = z - cz[a];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      for (int f = fn_off[a]; f < fn_off[a + 1]; ++f) acc += coef[f] * exp(-alpha[f] * r);
    }
    if (ok) out[p] = acc;
    return;
  }
  // the points' bounding box (idle threads repeat the last point)
  float bx[6] = {(float)x, (float)y, (float)z, -(float)x, -(float)y, -(float)z};  // min, then -max
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const float v = (float)-wave_min(-(double)-bx[c]);  // wave maximum of -bx = -(wave minimum of bx)
    if (lane == 0) s_red[c][wv] = -v;
  }
  __syncthreads();
  float bmin[3], bmax[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float lo = s_red[c][0], nhi = s_red[c + 3][0];
    for (int w = 1; w < nw; ++w) {
      lo = fminf(lo, s_red[c][w]);
      nhi = fminf(nhi, s_red[c + 3][w]);
    }
    bmin[c] = lo;
    bmax[c] = -nhi;
  }
  auto dist_box = [&](int a, float &dmin, float &dmax) {
    const float ac[3] = {(float)cx[a], (float)cy[a], (float)cz[a]};
    float n2 = 0.0f, f2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float lo = bmin[c] - ac[c], hi = ac[c] - bmax[c];
      const float dn = fmaxf(0.0f, fmaxf(lo, hi));
      const float df = fmaxf(fabsf(ac[c] - bmin[c]), fabsf(ac[c] - bmax[c]));
      n2 += dn * dn;
      f2 += df * df;
    }
    dmin = __builtin_amdgcn_sqrtf(n2) * (1.0f - 1e-6f);
    dmax = __builtin_amdgcn_sqrtf(f2) * (1.0f + 1e-6f) + 1e-6f;
  };
  constexpr float log2e = 1.4426950408889634f;
  // L = max_a log2 lo_a
  float L = -__builtin_inff();
  for (int a = threadIdx.x; a < natom; a += blockDim.x) {
    float dmin, dmax;
    dist_box(a, dmin, dmax);
    for (int f = fn_off[a]; f < fn_off[a + 1]; ++f)
      L = fmaxf(L, __builtin_amdgcn_logf((float)coef[f]) - (float)alpha[f] * log2e * dmax);
  }
  {
    const float wl = (float)-wave_min(-(double)L);
    if (lane == 0) s_lo[wv] = wl;
  }
  __syncthreads();
  for (int w = 0; w < nw; ++w) L = fmaxf(L, s_lo[w]);
  const float thr = L - 64.0f - 2.0f;
  // the kept atoms, 256 candidates at a time, in atom order
  for (int a0 = 0; a0 < natom; a0 += blockDim.x) {
    const int a = a0 + (int)threadIdx.x;
    bool keep = false;
    if (a < natom) {
      float dmin, dmax;
      dist_box(a, dmin, dmax);
      const int f0 = fn_off[a], nf = fn_off[a + 1] - f0;
      float u = -__builtin_inff();
      for (int f = f0; f < f0 + nf; ++f)
        u = fmaxf(u, __builtin_amdgcn_logf((float)coef[f]) - (float)alpha[f] * log2e * dmin);
      keep = nf > 0 && !(u + __builtin_amdgcn_logf((float)nf) < thr);
    }
    const unsigned long long bal = __ballot(keep);
    if (lane == 0) s_cnt[wv] = __popcll(bal);
    __syncthreads();
    int base = 0, nk = 0;
    for (int w = 0; w < nw; ++w) {
      base += w < wv ? s_cnt[w] : 0;
      nk += s_cnt[w];
    }
    if (keep)
      s_keep[base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = a;
    __syncthreads();
    nk = __builtin_amdgcn_readfirstlane(nk);
    for (int i = 0; i < nk; ++i) {
      const int b = s_keep[i];
      const double dx = x - cx[b], dy = y - cy[b], dz = z - cz[b];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      for (int f = fn_off[b]; f < fn_off[b + 1]; ++f) acc += coef[f] * exp(-alpha[f] * r);
    }
    __syncthreads();  // s_keep and s_cnt are rewritten by the next round
  }
  if (ok) out[p] = acc;
}

// ---------------------------------------------------------------------------
// Batched outer loop: many molecules resident at once, each with its own
// HpSlot.  One tick = k_batch_begin, the weights launches, the per-atom
// launches and k_batch_finalize over the slots in flight; a molecule whose
// run is HP_RUN_IDLE is skipped by every kernel.
// ---------------------------------------------------------------------------

// k_iter_begin for every molecule in flight: one workgroup per slot.
static __global__ __launch_bounds__(256) void k_batch_begin(HpSlot *__restrict__ slots, const int *__restrict__ slot_list,
                                                     const double *__restrict__ par_m, double *__restrict__ old_m,
                                                     const double *__restrict__ par_l, double *__restrict__ old_l,
                                                     const double *__restrict__ norm,
                                                     const double *__restrict__ alpha_fixed, double *__restrict__ coef,
                                                     double *__restrict__ alpha) {
  HpSlot &sl = slots[slot_list[blockIdx.x]];
  if (sl.run == HP_RUN_IDLE || sl.phase) return;  // idle, or inner solves still running
  __syncthreads();  // every thread has read run before thread 0 writes the slot
  if (threadIdx.x == 0)
    sl.refresh = sl.run == HP_RUN_FP32 && sl.corr_period > 0 && sl.niter % sl.corr_period == 0;
  for (int i = threadIdx.x; i < sl.nf; i += blockDim.x) {
    const int f = sl.f0 + i;
    if (sl.scheme == 0) {
      const double N = par_m[2 * f], S = par_m[2 * f + 1];
      old_m[2 * f] = N;
      old_m[2 * f + 1] = S;
      coef[f] = N * (S * S * S) / (8.0 * HP_PI);
      alpha[f] = S;
    } else {
      old_l[f] = par_l[f];
      coef[f] = par_l[f] * norm[f];
      alpha[f] = alpha_fixed[f];
    }
  }
}

// k_finalize for every molecule in flight, plus the stage machine of the
// mixed path: FP32 until change < max(threshold, switch level), then FP64
// (at least n_polish iterations) until change < threshold; maxiter caps the
// total unless the polish needs more.
static __global__ __launch_bounds__(256) void k_batch_finalize(HpSlot *__restrict__ slots, const int *__restrict__ slot_list,
                                                        const double *__restrict__ ent_part,
                                                        const double *__restrict__ change_part,
                                                        const int *__restrict__ status, int *__restrict__ status_all,
                                                        double *__restrict__ par_m, const double *__restrict__ old_m,
                                                        double *__restrict__ par_l, const double *__restrict__ old_l,
                                                        const int *__restrict__ ist_it) {
  __shared__ double lds[32];
  __shared__ int s_rollback;
  HpSlot &sl = slots[slot_list[blockIdx.x]];
  if (sl.run == HP_RUN_IDLE) return;
  const int s0 = sl.s0, ns = sl.ns, a0 = sl.a0, na = sl.na, f0 = sl.f0, nf = sl.nf, scheme = sl.scheme;
  // an atom whose inner solve is still running keeps the molecule in this
  // outer iteration (phase 1: no begin, no weights pass until it ends)
  int pend = 0;
  for (int i = threadIdx.x; i < na; i += blockDim.x) pend |= ist_it[a0 + i] >= 0;
  pend = __syncthreads_or(pend);
  if (threadIdx.x == 0) sl.phase = pend ? 1 : 0;
  if (pend) return;
  double e = 0.0, dummy = 0.0;
  for (int s = threadIdx.x; s < ns; s += blockDim.x) e += ent_part[s0 + s];
  block_sum2(e, dummy, lds);
  for (int i = threadIdx.x; i < na; i += blockDim.x) status_all[a0 + i] |= status[a0 + i];
  if (threadIdx.x == 0) {
    double msd = 0.0;
    int err = 0, any = 0;
    for (int a = 0; a < na; ++a) {
      msd += change_part[a0 + a];
      err |= status[a0 + a] & HP_ST_FATAL;
      any |= status[a0 + a];
    }
    const double ch = sqrt(msd);
    const int n = sl.niter + 1;
    sl.niter = n;
    sl.change = ch;
    sl.entropy = e;
    sl.err = err;
    sl.any |= any;
    s_rollback = 0;
    if (err != 0) {
      sl.event = HP_EV_ERROR;
      sl.run = HP_RUN_IDLE;
    } else if (n >= sl.max_iter || (n >= sl.min_stop && ch < sl.thr)) {
      if (sl.run == HP_RUN_FP32) {
        // end of the FP32 stage: this iteration is redone in FP64, so the
        // FP64 stage continues from the parameters it started from
        s_rollback = 1;
        const int m1 = n - 1;
        const int np = max(1, sl.n_polish);
        sl.niter = m1;
        sl.min_stop = m1 + np;
        sl.max_iter = max(sl.maxiter, m1 + np);
        sl.thr = sl.threshold;
        sl.event = HP_EV_SWITCH;
      } else {
        sl.event = (n >= sl.min_stop && ch < sl.thr) ? HP_EV_CONVERGED : HP_EV_MAXITER;
      }
      sl.run = HP_RUN_IDLE;
    }
  }
  __syncthreads();
  if (s_rollback) {
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
      const int f = f0 + i;
      if (scheme == 0) {
        par_m[2 * f] = old_m[2 * f];
        par_m[2 * f + 1] = old_m[2 * f + 1];
      } else {
        par_l[f] = old_l[f];
      }
    }
  }
}

// Host (set_run): the host moves molecules between stages by value.
struct HpRunSet {
  int n;
  int run;
  int slot[62];
};
static __global__ void k_batch_set_run(HpSlot *__restrict__ slots, HpRunSet rs) {
  const int i = threadIdx.x;
  if (i < rs.n) {
    slots[rs.slot[i]].run = rs.run;
    slots[rs.slot[i]].event = HP_EV_RUNNING;
  }
}

// Admission: interleaved points (x, y, z) into the SoA arena columns.
static __global__ void k_batch_points(const double *__restrict__ aos, int64_t n, double *__restrict__ px,
                               double *__restrict__ py, double *__restrict__ pz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    px[i] = aos[3 * i];
    py[i] = aos[3 * i + 1];
    pz[i] = aos[3 * i + 2];
  }
}

// Admission: a molecule's local indices shifted to its arena ranges, per-atom
// shell and QP offsets, and its per-atom state cleared.  One workgroup.
struct HpAdmit {
  int a0, na, s0, ns, f0;
  int64_t n0, q0;
  int has_qp;
};
static __global__ __launch_bounds__(256) void k_batch_admit(HpAdmit ad, const double *__restrict__ coords,
                                                     const int *__restrict__ shell_atom_l,
                                                     const int64_t *__restrict__ shell_off_l,
                                                     const int *__restrict__ fn_off_l, double *__restrict__ cx,
                                                     double *__restrict__ cy, double *__restrict__ cz,
                                                     int *__restrict__ shell_atom, int64_t *__restrict__ shell_off,
                                                     int *__restrict__ atom_shell_off, int *__restrict__ fn_off,
                                                     int64_t *__restrict__ qp_off, int *__restrict__ inner_niter,
                                                     int *__restrict__ status, int *__restrict__ status_all,
                                                     double *__restrict__ charges, double *__restrict__ change_part,
                                                     int *__restrict__ ist_it) {
  const int t = threadIdx.x;
  for (int a = t; a < ad.na; a += blockDim.x) {
    ist_it[ad.a0 + a] = -1;  // no inner solve in progress
    cx[ad.a0 + a] = coords[3 * a];
    cy[ad.a0 + a] = coords[3 * a + 1];
    cz[ad.a0 + a] = coords[3 * a + 2];
    inner_niter[ad.a0 + a] = 0;
    status[ad.a0 + a] = 0;
    status_all[ad.a0 + a] = 0;
    charges[ad.a0 + a] = 0.0;
    change_part[ad.a0 + a] = 0.0;
  }
  for (int a = t; a <= ad.na; a += blockDim.x) fn_off[ad.a0 + a] = ad.f0 + fn_off_l[a];
  for (int s = t; s < ad.ns; s += blockDim.x) {
    const int a = shell_atom_l[s];
    shell_atom[ad.s0 + s] = ad.a0 + a;
    shell_off[ad.s0 + s] = ad.n0 + shell_off_l[s];
    // first shell of its atom (shells are sorted by atom, each atom has one)
    if (s == 0 || shell_atom_l[s - 1] != a) atom_shell_off[ad.a0 + a] = ad.s0 + s;
  }
  if (t == 0) {
    shell_off[ad.s0 + ad.ns] = ad.n0 + shell_off_l[ad.ns];
    atom_shell_off[ad.a0 + ad.na] = ad.s0 + ad.ns;
    if (ad.has_qp) {
      int64_t q = ad.q0;
      for (int a = 0; a <= ad.na; ++a) {
        qp_off[ad.a0 + a] = q;
        if (a < ad.na) {
          const int64_t K = fn_off_l[a + 1] - fn_off_l[a];
          q += K * K;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Measurement probes: a plain copy (HBM bandwidth and counter calibration)
// and the throughput of the exponential the weights pass uses.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(256) void k_copy16(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                       int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

#define HP_EXP_RATE_CHAINS 8
// iters x 8 independent exponentials per thread, arguments in [-30, 0]; the
// FP64 form is the weights pass's (argument in units of ln2/64)
template <typename R>
__global__ __launch_bounds__(256) void k_exp_rate(int iters, double *__restrict__ out) {
  __shared__ double s_t64[64];
  if (threadIdx.x < 64) s_t64[threadIdx.x] = hp_exp2_64[threadIdx.x];
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  R x[HP_EXP_RATE_CHAINS], acc[HP_EXP_RATE_CHAINS];
#pragma unroll
  for (int c = 0; c < HP_EXP_RATE_CHAINS; ++c) {
    x[c] = (R)(-0.5 - 3.5 * c - 1e-3 * (t & 255));
    acc[c] = (R)0;
  }
  const R step = (R)(1.0 / 1024.0);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < HP_EXP_RATE_CHAINS; ++c) {
      acc[c] += neg_exp_scaled<R>((R)-1, x[c], s_t64);
      x[c] -= step;
    }
  }
  R s = (R)0;
#pragma unroll
  for (int c = 0; c < HP_EXP_RATE_CHAINS; ++c) s += acc[c];
  out[t] = (double)s;
}
