// grid_type 2: per-atom solvers over the whole molecular grid
// (gisa.py:257-279 _update_propars_atom_molgrids, mbis.py:169-174, and the
// LISA solvers of alisa.py:193-291 / 753-938 with P = N points).
//
// The AIM density of atom a at point p is rho(p) * clip(rho0_a,old(p) /
// promol(p), 0, 1): promol is stored once per outer iteration (N doubles)
// and rho0_a,old is recomputed from the outer iteration's parameters, so no
// natom x N weight array exists.  Each inner iteration is one or two passes
// over (point chunk x atom) workgroups followed by a per-atom update:
//   phase 0  evaluate at the current parameters: change against the previous
//            inner iterate, population, minimum of pro, and the scheme's sums
//            (MBIS: N_k and N_k<r>; sc: new coefficients; GISA: q; Newton:
//            gradient on the vector units and the Hessian B diag(w rho_a /
//            pro^2) B^T on the matrix cores)
//   phase 1  m-newton line search: min over points of pro(c + x_t delta) for
//            every x_t of linspace(tau, 0, n) in one pass
//   phase 2  outer change: sum_p w_p (rho0_a,new - rho0_a,old)^2
//   phase 3  done
#pragma once
#include <type_traits>
#include "hp_device.h"

typedef double hp_d4 __attribute__((ext_vector_type(4)));
typedef float hp_f4 __attribute__((ext_vector_type(4)));

#define HP_GT2_KM 8                        // functions per atom on this path
#define HP_GT2_NS (4 + 2 * HP_GT2_KM)      // summed values per pass
#define HP_GT2_NV (HP_GT2_NS + HP_GT2_KM * HP_GT2_KM)  // + Hessian
#define HP_GT2_NLIN 32                     // line-search candidates per pass
#define HP_GT2_CHUNK 4096                  // points per workgroup (multiple of 256)

struct Gt2Atom {
  int phase;
  int it;
  int status;
  int have_old;
  int t_ok;  // accepted line-search candidate, -1 if none pending
  int pad;
  double change;  // last inner change, then this atom's outer msd
  double pop;
  double c[16];      // current parameters (MBIS: N0,S0,N1,S1,...)
  double cprev[16];  // previous inner iterate
  double cold[16];   // parameters of the outer iteration (weights)
  double delta[16];  // Newton step
};

// Shell terms N S^3/(8 pi) exp(-S r) (MBIS) or basis values norm exp(-alpha r^n).
template <typename R>
__device__ __forceinline__ void gt2_terms(bool mbis, int K, const double *par, const double *norm,
                                          const double *alpha, const double *order, R r, R *T) {
#pragma unroll
  for (int k = 0; k < HP_GT2_KM; ++k) {
    T[k] = (R)0;
    if (k < K) {
      if (mbis) {
        const double N = par[2 * k], S = par[2 * k + 1];
        T[k] = (R)(N * (S * S * S) / (8.0 * HP_PI)) * hp_exp<R>(-(R)S * r);
      } else {
        const double n = order[k];
        const R arg = (n == 1.0) ? r : ((n == 2.0) ? r * r : (R)pow((double)r, n));
        T[k] = (R)norm[k] * hp_exp<R>(-(R)alpha[k] * arg);
      }
    }
  }
}

// Line-search candidate c + x delta, rounded after the product and after the
// sum (as numpy's propars + x * delta, alisa.py:846-847): the line search and
// the update that adopts its step compute the same coefficients bit for bit.
__device__ __forceinline__ double gt2_step(double c, double x, double delta) { return __dadd_rn(c, __dmul_rn(x, delta)); }

// Pro-atom value from its terms, summed in FP64 (the same arithmetic as the
// line search's candidates, whatever precision the terms are in).
template <typename R>
__device__ __forceinline__ double gt2_sum(bool mbis, int K, const R *T, const double *c) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < HP_GT2_KM; ++k)
    if (k < K) s += mbis ? (double)T[k] : (double)T[k] * c[k];
  return s;
}

// Phase 0 and 2 pass.  Partial layout per (atom, chunk): [0] change^2
// [1] pop [2] min pro [3] outer msd [4 .. NS) scheme sums, [NS .. NV) Hessian.
// KIND: 0 MBIS, 1 linear basis without Hessian (sc, GISA), 2 Newton.
template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_gt2_pass(const int *__restrict__ done, DevMol m,
                                                  const double *__restrict__ promol,
                                                  const Gt2Atom *__restrict__ st, const int *__restrict__ fn_off,
                                                  const double *__restrict__ fn_norm,
                                                  const double *__restrict__ fn_alpha,
                                                  const double *__restrict__ fn_order, int scheme, int solver,
                                                  double dc, double *__restrict__ partial) {
  constexpr int KM = HP_GT2_KM, NS = HP_GT2_NS, NV = HP_GT2_NV;
  // Newton: the basis terms that give the pro-atom (its negativity test, the
  // line search's candidates, the ratios) are FP64 in both precisions; R is
  // the precision of the MFMA inputs and accumulation of the Hessian
  using TR = typename std::conditional<KIND == 2, double, R>::type;
  __shared__ R s_g[4][KM][64];
  __shared__ R s_b[4][KM][64];
  __shared__ double s_red[4][NV];
  if (done && *done) return;
  const int a = blockIdx.y;
  const Gt2Atom &S = st[a];
  const int phase = S.phase;
  if (phase != 0 && phase != 2) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int f0 = fn_off[a], K = fn_off[a + 1] - f0;
  constexpr bool mbis = KIND == 0;
  const double *norm = fn_norm + f0, *alpha = fn_alpha + f0, *order = fn_order + f0;
  const double ax = m.cx[a], ay = m.cy[a], az = m.cz[a];
  const bool newton = KIND == 2 && phase == 0;
  const bool have_old = S.have_old != 0;
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  acc[2] = 1e300;
  hp_d4 h64 = {0.0, 0.0, 0.0, 0.0};
  hp_f4 h32 = {0.f, 0.f, 0.f, 0.f};
  const int64_t p_begin = (int64_t)blockIdx.x * HP_GT2_CHUNK;
  const int64_t p_end = min((int64_t)m.npoint, p_begin + HP_GT2_CHUNK);
  for (int64_t base = p_begin + wv * 64; base < p_begin + HP_GT2_CHUNK; base += 256) {
    const int64_t p = base + lane;
    const bool ok = p < p_end;
    const int64_t pp = ok ? p : p_begin;
    const double dx = m.px[pp] - ax, dy = m.py[pp] - ay, dz = m.pz[pp] - az;
    const TR r = (TR)sqrt(dx * dx + dy * dy + dz * dz);
    const double wmol = ok ? m.wmol[pp] : 0.0;
    const double rho = ok ? m.rho[pp] : 0.0;
    const double pm = promol[pp];
    TR T[KM], Told[KM];
    gt2_terms<TR>(mbis, K, S.cold, norm, alpha, order, r, Told);
    const double proold = gt2_sum<TR>(mbis, K, Told, S.cold);
    if (mbis) {
      gt2_terms<TR>(true, K, S.c, norm, alpha, order, r, T);
    } else {
#pragma unroll
      for (int k = 0; k < KM; ++k) T[k] = Told[k];  // fixed basis
    }
    const double pro = gt2_sum<TR>(mbis, K, T, S.c);
    if (phase == 2) {
      const double d = pro - (double)proold;
      acc[3] += wmol * d * d;
      continue;
    }
    // AIM density with the outer iteration's weights (np.clip semantics)
    const double rhoa = ok ? aim_weight((double)proold, pm) * rho : 0.0;
    acc[1] += wmol * rhoa;
    if (ok) acc[2] = fmin(acc[2], pro);
    if (have_old) {
      double prev;
      if (mbis) {
        TR Tp[KM];
        gt2_terms<TR>(true, K, S.cprev, norm, alpha, order, r, Tp);
        prev = gt2_sum<TR>(true, K, Tp, S.cprev);
      } else {
        prev = gt2_sum<TR>(false, K, T, S.cprev);
      }
      const double e = prev - pro;
      acc[0] += wmol * e * e;
    }
    const bool sick = (rhoa < dc) || (pro < dc);
    const double ratio = sick ? 0.0 : rhoa / pro;
    if (mbis) {
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const double wt = wmol * ((double)T[k] * ratio);
        acc[4 + 2 * k] += wt;
        acc[4 + 2 * k + 1] += wt * (double)r;
      }
    } else if (KIND == 1 && solver == 1) {
#pragma unroll
      for (int k = 0; k < KM; ++k) acc[4 + k] += wmol * (double)T[k] * rhoa;
    } else if (KIND == 1) {
#pragma unroll
      for (int k = 0; k < KM; ++k) acc[4 + k] += wmol * (((double)T[k] * S.c[k]) * ratio);
    } else if (KIND == 2) {
      // the Hessian's weight w rho_a / pro^2, one division per point
      const double hw = sick ? 0.0 : ratio / pro * wmol;
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const double integ = (double)T[k] * ratio;
        acc[4 + k] += integ * wmol;
        s_g[wv][k][lane] = (R)((double)T[k] * hw);
        s_b[wv][k][lane] = (R)T[k];
      }
    }
    if (newton) {
      __syncthreads();
      // 16 rank-4 updates: rows i < K of A are w g_i, columns j < K of B are
      // B_j, over the 64 points of this step (lane l: i = l%16, point 4q+l/16)
      const int i = lane & 15, kk = lane >> 4;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int pt = 4 * q + kk;
        const R av = i < K ? s_g[wv][i][pt] : (R)0;
        const R bv = i < K ? s_b[wv][i][pt] : (R)0;
        if (sizeof(R) == sizeof(double))
          h64 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av, (double)bv, h64, 0, 0, 0);
        else
          h32 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av, (float)bv, h32, 0, 0, 0);
      }
      __syncthreads();
    }
  }
  // slot 2 is a minimum (the pro-atom's smallest value), the others sums
  const double lane_min = acc[2];
  acc[2] = 0.0;
  wave_sum_n<NS>(acc);
  acc[2] = wave_min(lane_min);
  if (lane == 0) {
    for (int q = 0; q < NS; ++q) s_red[wv][q] = acc[q];
  }
  if (newton) {
    // output layout (measured on gfx950): lane l, register v holds
    // D[4v + l/16][l%16] for the FP64 16x16x4 MFMA and D[4(l/16) + v][l%16]
    // for the FP32 one
    const int j = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = sizeof(R) == sizeof(double) ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v;
      const double h = sizeof(R) == sizeof(double) ? h64[v] : (double)h32[v];
      if (i < KM && j < KM) s_red[wv][NS + i * KM + j] = h;
    }
  }
  __syncthreads();
  const int nq = newton ? NV : NS;
  if (threadIdx.x < nq) {
    const int q = threadIdx.x;
    double v = s_red[0][q];
    for (int w = 1; w < 4; ++w) v = (q == 2) ? fmin(v, s_red[w][q]) : v + s_red[w][q];
    partial[((int64_t)a * gridDim.x + blockIdx.x) * NV + q] = v;
  }
}

// Phase 1: for every line-search candidate x_t, whether any point of the
// chunk has pro(c + x_t delta) = sum_k B_k (c_k + x_t delta_k) not above the
// negative cutoff (alisa.py:836-850).  Each lane keeps a 32-bit mask (bit t
// set = candidate t fails somewhere); masks are OR-ed over lanes, waves and
// chunks, and the update takes the first candidate whose bit is clear.
//
// The candidate's value is linear in x: with P0 = sum_k B_k c_k and
// P1 = sum_k B_k delta_k, v_t = P0 + x_t P1 up to rounding.  The value the
// test is defined on is the FMA chain sum_k B_k cand_tk over the rounded
// candidates cand_tk = fl(c_k + fl(x_t delta_k)); it differs from
// fl(P0 + x_t P1) by less than (2 + 2 K + 1) u A (u = 2^-53, K <= 8 terms,
// A = sum_k |B_k| (|c_k| + xm |delta_k|), xm >= |x_t|), plus at most a few
// units of 2^-1074 where products fall below the normal range.  A candidate
// whose linear value is farther than 64 u A + 1e-300 from the cutoff is
// decided by it; the others run the chain itself, so every decision is the
// chain's.
__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

template <typename R>
__global__ __launch_bounds__(256) void k_gt2_linesearch(const int *__restrict__ done, DevMol m,
                                                        const Gt2Atom *__restrict__ st,
                                                        const int *__restrict__ fn_off,
                                                        const double *__restrict__ fn_norm,
                                                        const double *__restrict__ fn_alpha,
                                                        const double *__restrict__ fn_order, double tau, int nlin,
                                                        double neg_cutoff, double *__restrict__ partial) {
  constexpr int KM = HP_GT2_KM, NV = HP_GT2_NV, NL = HP_GT2_NLIN;
  __shared__ int s_mask[4];
  __shared__ double cand[NL][KM];  // c + x_t delta (alisa.py:838-840)
  __shared__ double s_x[NL];       // x_t
  if (done && *done) return;
  const int a = blockIdx.y;
  const Gt2Atom &S = st[a];
  if (S.phase != 1) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int f0 = fn_off[a], K = fn_off[a + 1] - f0;
  const double *norm = fn_norm + f0, *alpha = fn_alpha + f0, *order = fn_order + f0;
  const double ax = m.cx[a], ay = m.cy[a], az = m.cz[a];
  if (threadIdx.x < NL * KM) {
    const int t = threadIdx.x / KM, k = threadIdx.x % KM;
    const double step = (0.0 - tau) / (double)(nlin - 1);
    const double x = (t == nlin - 1) ? 0.0 : (double)t * step + tau;
    cand[t][k] = gt2_step(S.c[k], x, S.delta[k]);
    if (k == 0) s_x[t] = x;
  }
  // the linear form's coefficients and its error scale
  const double xm = fabs(tau) > 1.0 ? fabs(tau) : 1.0;
  double c[KM], d[KM], wgt[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    c[k] = k < K ? S.c[k] : 0.0;
    d[k] = k < K ? S.delta[k] : 0.0;
    wgt[k] = fabs(c[k]) + xm * fabs(d[k]);
  }
  __syncthreads();
  int bad = 0;
  const int64_t p_begin = (int64_t)blockIdx.x * HP_GT2_CHUNK;
  const int64_t p_end = min((int64_t)m.npoint, p_begin + HP_GT2_CHUNK);
  for (int64_t p = p_begin + threadIdx.x; p < p_end; p += 256) {
    const double dx = m.px[p] - ax, dy = m.py[p] - ay, dz = m.pz[p] - az;
    const R r = (R)sqrt(dx * dx + dy * dy + dz * dz);
    R T[KM];
    gt2_terms<R>(false, K, nullptr, norm, alpha, order, r, T);
    double P0 = 0.0, P1 = 0.0, A = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      if (k < K) {
        const double tk = (double)T[k];
        P0 += tk * c[k];
        P1 += tk * d[k];
        A += fabs(tk) * wgt[k];
      }
    }
    const double e = 64.0 * 0x1p-53 * A + 1e-300;  // + the rounding of products in the subnormal range
    for (int t = 0; t < nlin; ++t) {
      if (bad & (1 << t)) continue;
      const double va = P0 + s_x[t] * P1;
      if (va - e > neg_cutoff) continue;  // above the cutoff in the chain too
      bool fails = true;                  // at or below it in the chain too
      if (!(va + e <= neg_cutoff)) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < KM; ++k)
          if (k < K) v += (double)T[k] * cand[t][k];
        fails = !(v > neg_cutoff);
      }
      if (fails) bad |= 1 << t;
    }
  }
  bad = wave_or(bad);
  if (lane == 0) s_mask[wv] = bad;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int mask = s_mask[0] | s_mask[1] | s_mask[2] | s_mask[3];
    partial[((int64_t)a * gridDim.x + blockIdx.x) * NV + 4] = (double)(unsigned)mask;
  }
}

// Per-atom update after a pass: reduce the chunks (fixed order, one value per
// thread) and advance the solver in thread 0.
__global__ __launch_bounds__(128) void k_gt2_update(const int *__restrict__ done, int nchunk,
                                                    Gt2Atom *__restrict__ st, const int *__restrict__ fn_off,
                                                    const double *__restrict__ qpP, const int64_t *__restrict__ qp_off,
                                                    const double *__restrict__ partial, int scheme, int solver,
                                                    double thr, int maxit, double neg_cutoff, double pop_cutoff,
                                                    double tau, int nlin, int *__restrict__ active,
                                                    int *__restrict__ nhess) {
  constexpr int KM = HP_GT2_KM, NS = HP_GT2_NS, NV = HP_GT2_NV;
  constexpr int UB = 8;  // chunk values loaded together before they are added in order
  __shared__ double v[NV];
  __shared__ double M[(HP_KMAX + 1) * (HP_KMAX + 2)];
  __shared__ int s_or[4];
  if (done && *done) return;
  const int a = blockIdx.x;
  Gt2Atom &S = st[a];
  const int phase = S.phase;
  if (phase == 3) return;
  const bool newton = scheme != 0 && (solver == 3 || solver == 4);
  if (phase == 1) {
    // the candidates' failure masks: an OR, in any order, over every thread
    unsigned mask = 0u;
    for (int c = threadIdx.x; c < nchunk; c += blockDim.x)
      mask |= (unsigned)partial[((int64_t)a * nchunk + c) * NV + 4];
    const int wm = wave_or((int)mask);
    if ((threadIdx.x & 63) == 0) s_or[threadIdx.x >> 6] = wm;
  } else if (threadIdx.x < NV) {
    // sums (and the minimum in slot 2) over the chunks in chunk order
    const int q = threadIdx.x;
    const bool is_min = q == 2;
    double acc = is_min ? 1e300 : 0.0;
    if (q < NS || (newton && phase == 0)) {
      const double *col = partial + (int64_t)a * nchunk * NV + q;
      int c = 0;
      for (; c + UB <= nchunk; c += UB) {
        double pv[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) pv[u] = col[(int64_t)(c + u) * NV];
#pragma unroll
        for (int u = 0; u < UB; ++u) acc = is_min ? fmin(acc, pv[u]) : acc + pv[u];
      }
      for (; c < nchunk; ++c) {
        const double pv = col[(int64_t)c * NV];
        acc = is_min ? fmin(acc, pv) : acc + pv;
      }
    }
    v[q] = acc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (phase == 1) {
    unsigned mask = 0u;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mask |= (unsigned)s_or[w];
    v[4] = (double)mask;
  }
  const int K = fn_off[a + 1] - fn_off[a];
  if (newton && phase == 0) nhess[a] += 1;  // this round's pass formed the Hessian
  bool keep = true;
  if (phase == 2) {
    S.change = v[3];
    S.phase = 3;
    keep = false;
  } else if (phase == 1) {
    const unsigned mask = (unsigned)v[4];
    int t_ok = -1;
    for (int t = 0; t < nlin; ++t)
      if (!((mask >> t) & 1u)) {
        t_ok = t;
        break;
      }
    if (t_ok < 0) {
      S.status |= HP_ST_LINESEARCH;
      S.phase = 3;
      keep = false;
    } else {
      const double step = (0.0 - tau) / (double)(nlin - 1);
      const double x = (t_ok == nlin - 1) ? 0.0 : (double)t_ok * step + tau;
      for (int k = 0; k < K; ++k) {
        S.cprev[k] = S.c[k];
        S.c[k] = gt2_step(S.c[k], x, S.delta[k]);
      }
      S.have_old = 1;
      S.phase = 0;
    }
  } else {
    S.pop = v[1];
    const double change = S.have_old ? sqrt(v[0]) : 1e100;
    const bool converged = change < thr;
    if (scheme == 1 && solver == 1) {
      double q[HP_KMAX], x[HP_KMAX];
      for (int k = 0; k < K; ++k) q[k] = -2.0 * v[4 + k];
      if (!qp_simplex(qpP + qp_off[a], q, S.pop, K, x, M)) S.status |= HP_ST_QPFAIL;
      double sx = 0.0;
      for (int k = 0; k < K; ++k) {
        if (x[k] < neg_cutoff) S.status |= HP_ST_NEGPARS;
        sx += x[k];
        S.c[k] = x[k];
      }
      if (fabs(sx - S.pop) > pop_cutoff) S.status |= HP_ST_POP;
      S.it = 1;
      S.phase = 2;
    } else if (!newton) {
      // MBIS / sc: update first, then the stop test (mbis.py:140-160)
      const int np = scheme == 0 ? 2 * K : K;
      for (int k = 0; k < np; ++k) S.cprev[k] = S.c[k];
      if (scheme == 0) {
        for (int k = 0; k < K; ++k) {
          S.c[2 * k] = v[4 + 2 * k];
          S.c[2 * k + 1] = 3.0 * v[4 + 2 * k] / v[4 + 2 * k + 1];
        }
      } else {
        for (int k = 0; k < K; ++k) S.c[k] = v[4 + k];
      }
      S.it += 1;
      S.have_old = 1;
      if (converged) {
        if (scheme == 0) {
          double sumN = 0.0;
          for (int k = 0; k < K; ++k) sumN += S.c[2 * k];
          if (!(fabs(S.pop - sumN) <= 1e-4 + 1e-5 * fabs(sumN))) S.status |= HP_ST_POP;
        } else {
          double sx = 0.0;
          for (int k = 0; k < K; ++k) sx += S.c[k];
          if (fabs(sx - S.pop) > pop_cutoff) S.status |= HP_ST_POP;
        }
        S.phase = 2;
      } else if (S.it >= maxit) {
        S.status |= HP_ST_NOCONV;
        S.phase = 2;
      }
    } else {
      // Newton (alisa.py:884-938): the loop makes at most maxit evaluations
      if (S.it >= maxit) {
        S.status |= HP_ST_NEWTONMAX;
        S.phase = 3;
        keep = false;
      } else if (converged) {
        if (v[2] < neg_cutoff) S.status |= HP_ST_NEGDENS;
        double sx = 0.0;
        for (int k = 0; k < K; ++k) sx += S.c[k];
        if (fabs(sx - S.pop) > pop_cutoff) S.status |= HP_ST_POP;
        S.it += 1;
        S.phase = 2;
      } else {
        bool finite = true;
        for (int q = 0; q < K; ++q) finite = finite && isfinite(v[4 + q]);
        for (int i = 0; i < K; ++i)
          for (int j = 0; j < K; ++j) finite = finite && isfinite(v[NS + i * KM + j]);
        if (!finite) {
          S.status |= HP_ST_NONFINITE;
          S.phase = 3;
          keep = false;
        } else {
          const int ld = HP_KMAX + 2;
          for (int i = 0; i < K; ++i) {
            for (int j = 0; j < K; ++j) M[i * ld + j] = v[NS + i * KM + j];
            M[i * ld + K] = -1.0 - (-v[4 + i]);
          }
          double x[HP_KMAX];
          if (!dense_solve(M, K, ld, x)) {
            S.status |= HP_ST_SINGULAR;
            S.phase = 3;
            keep = false;
          } else {
            for (int k = 0; k < K; ++k) S.delta[k] = x[k];
            S.it += 1;
            if (solver == 3) {
              for (int k = 0; k < K; ++k) {
                S.cprev[k] = S.c[k];
                S.c[k] = S.c[k] + x[k];
              }
              S.have_old = 1;
            } else {
              S.phase = 1;
            }
          }
        }
      }
    }
  }
  if (keep) atomicAdd(active, 1);
}

// Start of the inner solves of an outer iteration.
__global__ void k_gt2_init(const int *__restrict__ done, int natom, Gt2Atom *__restrict__ st,
                           const int *__restrict__ fn_off, int scheme, const double *__restrict__ par) {
  if (done && *done) return;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natom) return;
  Gt2Atom &S = st[a];
  const int K = fn_off[a + 1] - fn_off[a];
  const int np = scheme == 0 ? 2 * K : K;
  const int off = scheme == 0 ? 2 * fn_off[a] : fn_off[a];
  S.phase = 0;
  S.it = 0;
  S.status = 0;
  S.have_old = 0;
  S.t_ok = -1;
  S.change = 0.0;
  S.pop = 0.0;
  for (int k = 0; k < 16; ++k) {
    const double v = k < np ? par[off + k] : 0.0;
    S.c[k] = v;
    S.cprev[k] = v;
    S.cold[k] = v;
    S.delta[k] = 0.0;
  }
}

// End of the inner solves: parameters, charges, change parts, status.
__global__ void k_gt2_finish(const int *__restrict__ done, DevMol m, int natom, const Gt2Atom *__restrict__ st,
                             const int *__restrict__ fn_off, int scheme, double *__restrict__ par,
                             int *__restrict__ inner_niter, int *__restrict__ status, double *__restrict__ charges,
                             double *__restrict__ change_part) {
  if (done && *done) return;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natom) return;
  const Gt2Atom &S = st[a];
  const int K = fn_off[a + 1] - fn_off[a];
  const int np = scheme == 0 ? 2 * K : K;
  const int off = scheme == 0 ? 2 * fn_off[a] : fn_off[a];
  for (int k = 0; k < np; ++k) par[off + k] = S.c[k];
  inner_niter[a] = S.it;
  status[a] = S.status;
  charges[a] = m.pseudo[a] - S.pop;
  change_part[a] = S.change;
}

// Clear the active-atom counter.
__global__ void k_gt2_zero(int *__restrict__ active) { *active = 0; }
