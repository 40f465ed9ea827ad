// Host runtime and C ABI of the MI355X stockholder-partitioning library.
//
// The outer fixed-point loop of AbstractISAWPart.do_partitioning
// (core/iterstock.py:159-193) runs here, natively: per iteration one launch of
// the fused weights pass, one launch of the per-atom solver and one
// finalisation launch, then a 32-byte read-back of (change, entropy, status)
// that decides convergence exactly as iterstock.py:187 does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hortonpart.h"
#include "hp_device.h"
#include "hp_glisa.h"
#include "hp_gt2.h"
#include "hp_host.h"
#include "hp_table.h"

namespace {

using hp::dalloc;
using hp::fail;
using hp::g_err;
using hp::pick_ppl;

template <typename T>
int dupload(T **p, const T *host, size_t n, hipStream_t s) {
  int rc = dalloc(p, n);
  if (rc) return rc;
  if (n) HP_HIP(hipMemcpyAsync(*p, host, n * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

}  // namespace

struct hp_part {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {};
  int natom = 0, nshell = 0, nfn = 0, npar = 0;
  int64_t npoint = 0;
  int scheme = 0, solver = 0, grid_type = 1, fnkind = HP_FN_SLATER;
  int only_molgrid = 0;  // grid_type 3: moments on the molecular grid (the loop runs as grid_type 2)
  int weights_ready = 0; // a weights pass has run since the parameters were first set
  int maxL = 0, maxR = 0, maxK = 0;
  hp::WeightsShape wshape{1, 64, 1};  // weights pass: points per thread, threads, shells per workgroup
  int ngroup = 0;                     // weights workgroups
  int *grp_first = nullptr;           // device: first radial shell of each weights workgroup
  double threshold = 1e-6, inner_threshold = 1e-8, dc = 1e-15, neg_cutoff = -1e-12, pop_cutoff = 1e-4;
  double rc = 0.0, tau = 1.0;
  int maxiter = 500, inner_maxiter = 2000, nlin = 20, check_mono = 0;
  std::vector<int64_t> h_shell_off;
  std::vector<int> h_atom_shell_off;
  std::vector<double> h_order;  // NLIS: the fixed power n of each function
  // device: molecule
  double *cx = nullptr, *cy = nullptr, *cz = nullptr, *px = nullptr, *py = nullptr, *pz = nullptr;
  double *rho = nullptr, *wmol = nullptr, *wat = nullptr, *wsph = nullptr, *shell_r = nullptr, *shell_w = nullptr,
         *pseudo = nullptr;
  int *shell_atom = nullptr, *atom_shell_off = nullptr;
  int64_t *shell_off = nullptr;
  // device: scheme
  int *fn_off = nullptr, *nb_list = nullptr, *shell_nb = nullptr;
  double *fn_norm = nullptr, *fn_alpha = nullptr, *fn_order = nullptr, *coef = nullptr, *alpha = nullptr,
         *qpP = nullptr;
  int64_t *qp_off = nullptr;
  // device: state
  double *par = nullptr, *old_par = nullptr, *rhobar = nullptr, *ent_part = nullptr, *promol = nullptr,
         *charges = nullptr, *change_part = nullptr, *out = nullptr, *mom_part = nullptr, *mom_out = nullptr,
         *scratch_pts = nullptr;
  size_t mom_part_cap = 0, mom_out_cap = 0;  // doubles held by mom_part, mom_out
  int *inner_niter = nullptr, *status = nullptr, *status_all = nullptr;
  // grid_type 2: per-atom solver states, chunk partials, active-atom counter
  Gt2Atom *gt2 = nullptr;
  double *gt2_partial = nullptr;
  int *gt2_active = nullptr;
  int *gt2_nhess = nullptr;  // per atom: Newton passes that formed a Hessian
  double evals_per_pass = 0.0;  // pro-atom function evaluations of one weights pass
  int gt2_nchunk = 0;
  double *hpar = nullptr, *hch = nullptr, *hent = nullptr, *hchg = nullptr;
  int hist_cap = 0;
  double *h_out = nullptr;  // pinned: change, entropy, error bits, any bits, done, iterations
  int *state = nullptr;     // device: done flag, iteration counter
  static constexpr int kSlots = 32;
  hipEvent_t slot_ev[kSlots][3] = {};
  double mixed_switch = 0.0;  // > 0: leave the FP32 stage below this change
  int mixed_refresh = 0;      // FP64 refresh of the FP32 correction every this many iterations
  int tail_cull = 1;          // weights pass may skip atoms bounded below one rounding
  int params_nonneg = 1;      // the parameters last set are all >= 0 (tail skipping needs it)
  unsigned long long *evals_done = nullptr;  // [FP32, FP64] evaluations of the last run
  // gLISA global terms (allocated at the first call)
  int gl_M = 0, gl_Mp = 0, gl_npair = 0, gl_nchunk = 0, gl_hnchunk = 0, gl_nblock = 0;
  double *gl_G = nullptr, *gl_D = nullptr, *gl_E = nullptr, *gl_obj = nullptr, *gl_grad = nullptr,
         *gl_hess = nullptr, *gl_out = nullptr, *gl_coef = nullptr, *gl_pro = nullptr;
  int2 *gl_pairs = nullptr;
  double *corr = nullptr;     // per radial shell: FP64 - FP32 spherical average
  // tabulated pro-atoms (ISA, HI, Hirshfeld): knots, their radial weights,
  // the current cubic pieces, HI's current knot densities and records, and
  // the per-shell atom-grid populations of the last weights pass
  int ntab = 0;
  int *tab_off = nullptr, *rec_off = nullptr, *rec_charge = nullptr, *tab_guide = nullptr;
  float2 *tab_gparam = nullptr;
  int64_t *rec_val = nullptr;
  double *tab_x = nullptr, *tab_w = nullptr, *tab_c = nullptr, *cur_rho = nullptr, *rec_rho = nullptr,
         *rec_c = nullptr, *pop_part = nullptr;
  // host state
  int niter = 0;
  double last_change = 0.0, last_entropy = 0.0;
  double k_ms[4] = {0, 0, 0, 0};
  int64_t k_n[4] = {0, 0, 0, 0};

  DevMol mol() const {
    DevMol m;
    m.natom = natom;
    m.nshell = nshell;
    m.npoint = npoint;
    m.cx = cx;
    m.cy = cy;
    m.cz = cz;
    m.px = px;
    m.py = py;
    m.pz = pz;
    m.rho = rho;
    m.wmol = wmol;
    m.wat = wat;
    m.wsph = wsph;
    m.shell_atom = shell_atom;
    m.shell_off = shell_off;
    m.shell_r = shell_r;
    m.shell_w = shell_w;
    m.atom_shell_off = atom_shell_off;
    m.pseudo = pseudo;
    m.work = nullptr;
    m.slots = nullptr;
    m.grp_first = grp_first;
    m.wgroup = wshape.g;
    return m;
  }
  DevFn fn() const {
    DevFn f;
    f.fn_off = fn_off;
    f.coef = coef;
    f.alpha = alpha;
    f.order = fn_order;
    f.nb_list = rc > 0.0 ? nb_list : nullptr;
    f.shell_nb = shell_nb;
    f.nb_off = nullptr;
    f.rc = rc;
    // MBIS and GISA updates keep non-negative coefficients non-negative
    f.tail_cull = scheme != HP_SCHEME_LISA && fnkind != HP_FN_GENERAL && tail_cull && params_nonneg;
    f.evals_done = evals_done;
    f.tab_off = tab_off;
    f.tab_x = tab_x;
    f.tab_c = tab_c;
    f.tab_guide = tab_guide;
    f.tab_gparam = tab_gparam;
    return f;
  }
  DevRec rec() const {
    DevRec r;
    r.rec_off = rec_off;
    r.rec_charge = rec_charge;
    r.rec_val = rec_val;
    r.rec_rho = rec_rho;
    r.rec_c = rec_c;
    return r;
  }
  bool tabulated() const {
    return scheme == HP_SCHEME_ISA || scheme == HP_SCHEME_HI || scheme == HP_SCHEME_HIRSHFELD;
  }
};

namespace {

// k_iter_begin's parameter layout: 0 MBIS (N, S), 1 linear coefficients,
// 2 NLIS/GMBIS (N, S, n)
int table_kind(const hp_part *p) {
  return p->scheme == HP_SCHEME_MBIS ? 0 : (p->scheme == HP_SCHEME_NLIS ? 2 : 1);
}

void free_part(hp_part *p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  void *bufs[] = {p->cx,      p->cy,          p->cz,      p->px,        p->py,          p->pz,       p->rho,
                  p->wmol,    p->wat,         p->wsph,    p->shell_r,   p->shell_w,     p->pseudo,   p->shell_atom,
                  p->atom_shell_off, p->shell_off, p->fn_off, p->nb_list, p->shell_nb,    p->fn_norm,  p->fn_alpha,
                  p->fn_order, p->coef,       p->alpha,   p->qpP,       p->qp_off,      p->par,      p->old_par,
                  p->rhobar,  p->ent_part,    p->promol,  p->charges,   p->change_part, p->out,      p->mom_part,
                  p->mom_out, p->scratch_pts, p->inner_niter, p->status, p->status_all, p->hpar,    p->hch,
                  p->hent,    p->hchg,    p->corr,      p->evals_done, p->grp_first, p->gl_G, p->gl_D,
                  p->gl_E,    p->gl_obj,  p->gl_grad,   p->gl_hess,  p->gl_out,    p->gl_coef, p->gl_pro,
                  p->gl_pairs, p->tab_off, p->rec_off, p->rec_charge, p->rec_val, p->tab_x, p->tab_w,
                  p->tab_c,   p->cur_rho, p->rec_rho,   p->rec_c,    p->pop_part, p->tab_guide, p->tab_gparam};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (p->h_out) (void)hipHostFree(p->h_out);
  if (p->state) (void)hipFree(p->state);
  if (p->gt2) (void)hipFree(p->gt2);
  if (p->gt2_partial) (void)hipFree(p->gt2_partial);
  if (p->gt2_active) (void)hipFree(p->gt2_active);
  if (p->gt2_nhess) (void)hipFree(p->gt2_nhess);
  for (auto &e : p->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &slot : p->slot_ev)
    for (auto &e : slot)
      if (e) (void)hipEventDestroy(e);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// ---- kernel launch helpers -------------------------------------------------

template <typename R, int FK, bool GT2, int PT, int KM>
void launch_weights_km(hp_part *p, int bs, int tile, int corr_mode) {
  const size_t lds = hp::weights_lds(weights_lds_records<R>(KM, tile, FK), p->wshape);
  hipLaunchKernelGGL((k_weights<R, FK, GT2, PT, KM>), dim3(p->ngroup), dim3(bs), lds, p->stream, p->state, p->mol(),
                     p->fn(), p->rhobar, p->ent_part, p->promol, p->dc, tile, p->corr, corr_mode);
}

template <typename R, int FK, bool GT2, int PT>
void launch_weights_pt(hp_part *p, int bs, int corr_mode) {
  const int km = weights_km(p->maxK, FK);
  const int tile = weights_tile_size<R>(km, p->natom);
  if constexpr (FK == HP_FN_GENERAL) {
    launch_weights_km<R, FK, GT2, PT, 16>(p, bs, tile, corr_mode);
  } else {
    if (km == 2)
      launch_weights_km<R, FK, GT2, PT, 2>(p, bs, tile, corr_mode);
    else if (km == 4)
      launch_weights_km<R, FK, GT2, PT, 4>(p, bs, tile, corr_mode);
    else if (km == 8)
      launch_weights_km<R, FK, GT2, PT, 8>(p, bs, tile, corr_mode);
    else
      launch_weights_km<R, FK, GT2, PT, 16>(p, bs, tile, corr_mode);
  }
}

template <typename R, int FK>
void launch_weights_fk(hp_part *p, bool gt2, int corr_mode) {
  // one workgroup per group of radial shells, each thread one or two points
  const int pt = p->wshape.pt;
  const int bs = p->wshape.bs;
  if (gt2) {
    if (pt == 1)
      launch_weights_pt<R, FK, true, 1>(p, bs, HP_CORR_NONE);
    else
      launch_weights_pt<R, FK, true, 2>(p, bs, HP_CORR_NONE);
  } else {
    if (pt == 1)
      launch_weights_pt<R, FK, false, 1>(p, bs, corr_mode);
    else
      launch_weights_pt<R, FK, false, 2>(p, bs, corr_mode);
  }
}

template <typename R>
void launch_weights(hp_part *p, bool gt2, int corr_mode = HP_CORR_NONE) {
  if (p->fnkind == HP_FN_SLATER)
    launch_weights_fk<R, HP_FN_SLATER>(p, gt2, corr_mode);
  else if (p->fnkind == HP_FN_GAUSS)
    launch_weights_fk<R, HP_FN_GAUSS>(p, gt2, corr_mode);
  else
    launch_weights_fk<R, HP_FN_GENERAL>(p, gt2, corr_mode);
}

template <int PPL>
int launch_inner_ppl(hp_part *p) {
  const dim3 grid((p->natom + 3) / 4), block(256);
  if (p->scheme == HP_SCHEME_MBIS) {
    // one wavefront per atom with two to four radial points per lane (up to
    // 256 points), else one workgroup per atom with one point per lane
    const int ppl = hp::mbis_ppl(p->maxR);
    const dim3 g1(p->natom), b1(ppl > 1 ? 64 : ((p->maxR + 63) / 64) * 64);
#define HP_MBIS(KMV, PPLV)                                                                                            \
  hipLaunchKernelGGL((k_mbis_inner<KMV, PPLV>), g1, b1, 0, p->stream, p->state, p->mol(), p->rhobar, p->par,          \
                     p->old_par, p->fn_off, p->inner_threshold, p->inner_maxiter, p->inner_niter, p->status,          \
                     p->charges, p->change_part, 0, nullptr, nullptr)
#define HP_MBIS_K(KMV) \
  do {                    \
    if (ppl == 2)         \
      HP_MBIS(KMV, 2);    \
    else if (ppl == 3)    \
      HP_MBIS(KMV, 3);    \
    else if (ppl == 4)    \
      HP_MBIS(KMV, 4);    \
    else                  \
      HP_MBIS(KMV, 1);    \
  } while (0)
    if (p->maxK <= 2)
      HP_MBIS_K(2);
    else if (p->maxK <= 3)
      HP_MBIS_K(3);
    else
      HP_MBIS_K(7);
#undef HP_MBIS_K
#undef HP_MBIS
  } else if (p->scheme == HP_SCHEME_NLIS) {
    const dim3 g1(p->natom), b1(((p->maxR + 63) / 64) * 64);
    hipLaunchKernelGGL(k_nlis_inner, g1, b1, 0, p->stream, p->state, p->mol(), p->rhobar, p->par, p->old_par,
                       p->fn_off, p->fn_norm, p->inner_threshold, p->inner_maxiter, p->inner_niter, p->status,
                       p->charges, p->change_part);
  } else if (p->solver == HP_SOLVER_QP) {
    hipLaunchKernelGGL((k_gisa_qp<PPL>), grid, block, 0, p->stream, p->state, p->mol(), p->rhobar, p->par, p->old_par, p->fn_off,
                       p->fn_norm, p->fn_alpha, p->fn_order, p->qpP, p->qp_off, p->pop_cutoff, p->neg_cutoff,
                       p->inner_niter, p->status, p->charges, p->change_part);
  } else {
#define HP_LISA(KMV)                                                                                             \
  hipLaunchKernelGGL((k_lisa_inner<PPL, KMV>), grid, block, 0, p->stream, p->state, p->mol(), p->rhobar, p->par, p->old_par, \
                     p->fn_off, p->fn_norm, p->fn_alpha, p->fn_order, p->solver, p->inner_threshold,             \
                     p->inner_maxiter, p->dc, p->neg_cutoff, p->pop_cutoff, p->tau, p->nlin, p->check_mono,       \
                     p->inner_niter, p->status, p->charges, p->change_part)
    if (p->maxK <= 4)
      HP_LISA(4);
    else if (p->maxK <= 6)
      HP_LISA(6);
    else if (p->maxK <= 8)
      HP_LISA(8);
    else if (p->maxK <= 12)
      HP_LISA(12);
    else
      return fail(HP_E_UNSUPP, "LISA on the device supports at most 12 functions per atom");
#undef HP_LISA
  }
  return 0;
}

int launch_inner(hp_part *p) {
  const int ppl = pick_ppl(p->maxR);
  if (ppl == 2) return launch_inner_ppl<2>(p);
  if (ppl == 4) return launch_inner_ppl<4>(p);
  if (ppl == 8) return launch_inner_ppl<8>(p);
  return fail(HP_E_UNSUPP, "more than 512 radial points per atom");
}

// Enqueue one outer iteration (update_at_weights, per-atom updates, entropy,
// change, stop test).  Every kernel returns at once when the device flag says
// the loop has stopped, so iterations can be enqueued ahead of the decision.
int enqueue_iteration(hp_part *p, int precision, double thr, int min_stop, int max_iter, int slot, bool refresh,
                      int rollback) {
  hipStream_t s = p->stream;
  const int n = std::max(p->npar, p->nfn);
  hipLaunchKernelGGL(k_iter_begin, dim3((n + 255) / 256), dim3(256), 0, s, p->state, table_kind(p),
                     p->npar, p->nfn, p->par, p->old_par, p->fn_norm, p->fn_alpha, p->coef, p->alpha);
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->slot_ev[slot][0], s));
  if (precision == HP_PREC_MIXED) {
    if (p->mixed_refresh <= 0) {
      launch_weights<float>(p, false, HP_CORR_NONE);
    } else if (refresh) {
      // FP64 iteration that also calibrates the FP32 correction
      launch_weights<double>(p, false, HP_CORR_NONE);
      HP_LAUNCHED();
      launch_weights<float>(p, false, HP_CORR_CALIBRATE);
    } else {
      launch_weights<float>(p, false, HP_CORR_APPLY);
    }
  } else {
    launch_weights<double>(p, false);
  }
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->slot_ev[slot][1], s));
  int rc = launch_inner(p);
  if (rc) return rc;
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->slot_ev[slot][2], s));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, p->state, thr, min_stop, max_iter, p->nshell, p->ent_part,
                     p->natom, p->change_part, p->status, p->status_all, p->par, p->npar, p->charges, p->hpar, p->hch,
                     p->hent, p->hchg, p->out, p->old_par, rollback);
  HP_LAUNCHED();
  return 0;
}

// Build the tabulated pro-atoms from the parameters (ISA, HI; Hirshfeld's
// tables are fixed at creation).
void launch_table_begin(hp_part *p) {
  hipStream_t s = p->stream;
  if (p->scheme == HP_SCHEME_ISA)
    hipLaunchKernelGGL(k_isa_begin, dim3(p->natom), dim3(64), 0, s, p->state, p->tab_off, p->tab_x, p->par, p->old_par,
                       p->tab_c);
  else if (p->scheme == HP_SCHEME_HI)
    hipLaunchKernelGGL(k_hi_begin, dim3(p->natom), dim3(64), 0, s, p->state, p->natom, p->tab_off, p->rec(), p->pseudo,
                       p->par, p->old_par, p->tab_c, p->cur_rho, p->status);
}

void launch_table_weights(hp_part *p) {
  hipLaunchKernelGGL(k_table_weights, dim3(p->nshell), dim3(256), 0, p->stream, p->state, p->mol(), p->fn(), p->dc,
                     p->rhobar, p->pop_part, p->ent_part, p->promol);
}

// One outer iteration of ISA or HI: tables, weights pass, per-atom update,
// stop test (the kernels skip their work once the device flag is set).
int enqueue_iteration_table(hp_part *p, double thr, int min_stop, int max_iter, int slot) {
  hipStream_t s = p->stream;
  launch_table_begin(p);
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->slot_ev[slot][0], s));
  launch_table_weights(p);
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->slot_ev[slot][1], s));
  if (p->scheme == HP_SCHEME_ISA)
    hipLaunchKernelGGL(k_isa_update, dim3(p->natom), dim3(64), 0, s, p->state, p->mol(), p->tab_off, p->tab_x, p->tab_w,
                       p->rhobar, p->par, p->old_par, p->charges, p->change_part, p->status);
  else
    hipLaunchKernelGGL(k_hi_update, dim3(p->natom), dim3(64), 0, s, p->state, p->mol(), p->tab_off, p->tab_x, p->tab_w,
                       p->rec(), p->pop_part, p->cur_rho, p->par, p->charges, p->change_part, p->status);
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->slot_ev[slot][2], s));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, p->state, thr, min_stop, max_iter, p->nshell, p->ent_part,
                     p->natom, p->change_part, p->status, p->status_all, p->par, p->npar, p->charges, p->hpar, p->hch,
                     p->hent, p->hchg, p->out, p->old_par, 0);
  HP_LAUNCHED();
  return 0;
}

// Run iterations until the device stop test fires: batches of speculative
// iterations, one host synchronisation per batch.
int run_stage(hp_part *p, int precision, double thr, int min_stop, int max_iter, int rollback = 0) {
  hipStream_t s = p->stream;
  HP_HIP(hipMemsetAsync(p->state, 0, sizeof(int), s));
  int batch = 2;
  while (true) {
    const int before = p->niter;
    const int remaining = max_iter - before;
    const int nb = std::max(1, std::min(batch, remaining));
    for (int c = 0; c < nb; ++c) {
      // iteration before + c, if the loop has not stopped by then
      const bool refresh = p->mixed_refresh > 0 && (before + c) % p->mixed_refresh == 0;
      int rc = p->tabulated() ? enqueue_iteration_table(p, thr, min_stop, max_iter, c)
                              : enqueue_iteration(p, precision, thr, min_stop, max_iter, c, refresh, rollback);
      if (rc) return rc;
    }
    HP_HIP(hipMemcpyAsync(p->h_out, p->out, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    HP_HIP(hipMemcpyAsync(p->h_out + 4, p->state, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    int st[2];
    std::memcpy(st, p->h_out + 4, sizeof(st));
    p->niter = st[1];
    const int executed = p->niter - before;
    for (int c = 0; c < executed && c < nb; ++c) {
      float ms = 0.f;
      HP_HIP(hipEventElapsedTime(&ms, p->slot_ev[c][0], p->slot_ev[c][1]));
      p->k_ms[HP_KERNEL_WEIGHTS] += ms;
      p->k_n[HP_KERNEL_WEIGHTS] += 1;
      HP_HIP(hipEventElapsedTime(&ms, p->slot_ev[c][1], p->slot_ev[c][2]));
      p->k_ms[HP_KERNEL_INNER] += ms;
      p->k_n[HP_KERNEL_INNER] += 1;
    }
    p->last_change = p->h_out[0];
    p->last_entropy = p->h_out[1];
    if ((int)p->h_out[2] != 0) return fail(HP_E_SOLVER, "a per-atom solver raised; see hp_part_status");
    if (st[0]) break;
    batch = std::min(batch * 2, hp_part::kSlots);
  }
  return 0;
}

}  // namespace

namespace {

// One outer iteration on the molecular grid (grid_type 2): weights pass that
// stores the promolecule, then rounds of per-atom passes until every atom's
// solver has finished, then the finalisation.  Returns with the device state
// read back (done flag, iteration count).
int gt2_iteration(hp_part *p, int precision, double thr, int min_stop, int max_iter) {
  hipStream_t s = p->stream;
  const int n = std::max(p->npar, p->nfn);
  hipLaunchKernelGGL(k_iter_begin, dim3((n + 255) / 256), dim3(256), 0, s, p->state, table_kind(p),
                     p->npar, p->nfn, p->par, p->old_par, p->fn_norm, p->fn_alpha, p->coef, p->alpha);
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->ev[0], s));
  if (precision == HP_PREC_MIXED)
    launch_weights<float>(p, true);
  else
    launch_weights<double>(p, true);
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->ev[1], s));
  hipLaunchKernelGGL(k_gt2_init, dim3((p->natom + 127) / 128), dim3(128), 0, s, p->state, p->natom, p->gt2, p->fn_off,
                     p->scheme == HP_SCHEME_MBIS ? 0 : 1, p->par);
  HP_LAUNCHED();
  const double dc = p->scheme == HP_SCHEME_MBIS ? 1e-15 : p->dc;  // mbis.py:82 default cutoff
  const int scheme = p->scheme == HP_SCHEME_MBIS ? 0 : p->scheme;
  const dim3 grid(p->gt2_nchunk, p->natom), block(256);
  // every round advances each unfinished atom by one step; bound the rounds
  const int max_rounds = 2 * p->inner_maxiter + 8;
  int rounds = 0;
  int *h_active = reinterpret_cast<int *>(p->h_out + 6);
  bool finished = false;
  while (!finished) {
    const int nb = 2;
    for (int c = 0; c < nb; ++c) {
      hipLaunchKernelGGL(k_gt2_zero, dim3(1), dim3(1), 0, s, p->gt2_active);
      HP_HIP(hipEventRecord(p->slot_ev[c][0], s));
      const int kind = scheme == 0 ? 0 : ((p->solver == HP_SOLVER_NEWTON || p->solver == HP_SOLVER_MNEWTON) ? 2 : 1);
#define HP_GT2_PASS(RT, KD)                                                                                      \
  hipLaunchKernelGGL((k_gt2_pass<RT, KD>), grid, block, 0, s, p->state, p->mol(), p->promol, p->gt2, p->fn_off, \
                     p->fn_norm, p->fn_alpha, p->fn_order, scheme, p->solver, dc, p->gt2_partial)
      if (precision == HP_PREC_MIXED) {
        if (kind == 0)
          HP_GT2_PASS(float, 0);
        else if (kind == 1)
          HP_GT2_PASS(float, 1);
        else
          HP_GT2_PASS(float, 2);
      } else {
        if (kind == 0)
          HP_GT2_PASS(double, 0);
        else if (kind == 1)
          HP_GT2_PASS(double, 1);
        else
          HP_GT2_PASS(double, 2);
      }
#undef HP_GT2_PASS
      HP_LAUNCHED();
      HP_HIP(hipEventRecord(p->slot_ev[c][1], s));
      if (p->solver == HP_SOLVER_MNEWTON) {
        // FP64 in both precisions: it tests the same pro-atom values as the pass
        hipLaunchKernelGGL(k_gt2_linesearch<double>, grid, block, 0, s, p->state, p->mol(), p->gt2, p->fn_off,
                           p->fn_norm, p->fn_alpha, p->fn_order, p->tau, p->nlin, p->neg_cutoff, p->gt2_partial);
        HP_LAUNCHED();
      }
      hipLaunchKernelGGL(k_gt2_update, dim3(p->natom), dim3(128), 0, s, p->state, p->gt2_nchunk, p->gt2, p->fn_off,
                         p->qpP, p->qp_off, p->gt2_partial, scheme, p->solver, p->inner_threshold, p->inner_maxiter,
                         p->neg_cutoff, p->pop_cutoff, p->tau, p->nlin, p->gt2_active, p->gt2_nhess);
      HP_LAUNCHED();
      HP_HIP(hipEventRecord(p->slot_ev[c][2], s));
    }
    HP_HIP(hipMemcpyAsync(h_active, p->gt2_active, sizeof(int), hipMemcpyDeviceToHost, s));
    HP_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nb; ++c) {
      float ms = 0.f;
      HP_HIP(hipEventElapsedTime(&ms, p->slot_ev[c][0], p->slot_ev[c][1]));
      p->k_ms[HP_KERNEL_LISA_GRID] += ms;
      p->k_n[HP_KERNEL_LISA_GRID] += 1;
      HP_HIP(hipEventElapsedTime(&ms, p->slot_ev[c][0], p->slot_ev[c][2]));
      p->k_ms[HP_KERNEL_INNER] += ms;
      p->k_n[HP_KERNEL_INNER] += 1;
    }
    rounds += nb;
    finished = *h_active == 0;
    if (!finished && rounds >= max_rounds) return fail(HP_E_SOLVER, "grid_type 2 inner solver did not finish");
  }
  float ms = 0.f;
  HP_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
  p->k_ms[HP_KERNEL_WEIGHTS] += ms;
  p->k_n[HP_KERNEL_WEIGHTS] += 1;
  hipLaunchKernelGGL(k_gt2_finish, dim3((p->natom + 127) / 128), dim3(128), 0, s, p->state, p->mol(), p->natom, p->gt2,
                     p->fn_off, p->scheme == HP_SCHEME_MBIS ? 0 : 1, p->par, p->inner_niter, p->status, p->charges,
                     p->change_part);
  HP_LAUNCHED();
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, p->state, thr, min_stop, max_iter, p->nshell, p->ent_part,
                     p->natom, p->change_part, p->status, p->status_all, p->par, p->npar, p->charges, p->hpar, p->hch,
                     p->hent, p->hchg, p->out, p->old_par, 0);
  HP_LAUNCHED();
  HP_HIP(hipMemcpyAsync(p->h_out, p->out, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  HP_HIP(hipMemcpyAsync(p->h_out + 4, p->state, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  return 0;
}

// Partial sums of the moments: one row of HP_NMOM per radial shell (atom
// grids) or per (atom, point chunk) (grid_type 3, the molecular grid).
size_t moments_partials(const hp_part *p, bool molgrid) {
  if (!molgrid) return (size_t)p->nshell;
  return (size_t)p->natom * (size_t)((p->npoint + HP_MOM_CHUNK - 1) / HP_MOM_CHUNK);
}

// Launch the moments passes: partials into part_buf (moments_partials rows),
// per-atom sums into out_buf (natom rows).  grid_type 3: every atom's AIM
// density on the whole molecular grid with the molecular weights
// (core/base.py:352-402 with local = False); otherwise on the atom's own grid
// with the atom-grid weights.
int launch_moments(hp_part *p, double *part_buf, double *out_buf, bool molgrid) {
  DevFn f = p->fn();
  if (molgrid) {
    const int nchunk = (int)((p->npoint + HP_MOM_CHUNK - 1) / HP_MOM_CHUNK);
    const dim3 g((unsigned)nchunk, (unsigned)p->natom);
    if (p->fnkind == HP_FN_TABLE)
      hipLaunchKernelGGL((k_moments_mol<HP_FN_TABLE>), g, dim3(256), 0, p->stream, p->mol(), f, p->promol, part_buf);
    else if (p->fnkind == HP_FN_SLATER)
      hipLaunchKernelGGL((k_moments_mol<HP_FN_SLATER>), g, dim3(256), 0, p->stream, p->mol(), f, p->promol, part_buf);
    else if (p->fnkind == HP_FN_GAUSS)
      hipLaunchKernelGGL((k_moments_mol<HP_FN_GAUSS>), g, dim3(256), 0, p->stream, p->mol(), f, p->promol, part_buf);
    else
      hipLaunchKernelGGL((k_moments_mol<HP_FN_GENERAL>), g, dim3(256), 0, p->stream, p->mol(), f, p->promol, part_buf);
    HP_LAUNCHED();
    hipLaunchKernelGGL(k_moments_reduce_mol, dim3(p->natom), dim3(64), 0, p->stream, nchunk, part_buf, out_buf);
  } else {
    if (p->fnkind == HP_FN_TABLE)
      hipLaunchKernelGGL((k_moments<HP_FN_TABLE>), dim3(p->nshell), dim3(256), 0, p->stream, p->mol(), f, part_buf);
    else if (p->fnkind == HP_FN_SLATER)
      hipLaunchKernelGGL((k_moments<HP_FN_SLATER>), dim3(p->nshell), dim3(256), 0, p->stream, p->mol(), f, part_buf);
    else if (p->fnkind == HP_FN_GAUSS)
      hipLaunchKernelGGL((k_moments<HP_FN_GAUSS>), dim3(p->nshell), dim3(256), 0, p->stream, p->mol(), f, part_buf);
    else
      hipLaunchKernelGGL((k_moments<HP_FN_GENERAL>), dim3(p->nshell), dim3(256), 0, p->stream, p->mol(), f, part_buf);
    HP_LAUNCHED();
    hipLaunchKernelGGL(k_moments_reduce, dim3(p->natom), dim3(64), 0, p->stream, p->mol(), part_buf, out_buf);
  }
  HP_LAUNCHED();
  return 0;
}

// The moments of order lmax as rows of the per-atom sums: lmax <= 3 takes a
// prefix of each family from the 40 sums of k_moments (cart 20 | pure 16 |
// radial 4); higher orders have rows of nq = ncart + npure + nrad from
// k_moments_any.
struct MomShape {
  int ncart, npure, nrad;
  int stride, src_cart, src_pure, src_rad;
  bool fast;
};

MomShape moment_shape(int lmax) {
  MomShape s;
  s.ncart = (lmax + 1) * (lmax + 2) * (lmax + 3) / 6;
  s.npure = (lmax + 1) * (lmax + 1);
  s.nrad = lmax + 1;
  s.fast = lmax <= 3;
  if (s.fast) {
    s.stride = HP_NMOM;
    s.src_cart = 0;
    s.src_pure = 20;
    s.src_rad = 36;
  } else {
    s.stride = s.ncart + s.npure + s.nrad;
    s.src_cart = 0;
    s.src_pure = s.ncart;
    s.src_rad = s.ncart + s.npure;
  }
  return s;
}

// The moments passes for order lmax: partials into part_buf
// (moments_partials rows of ms.stride), per-atom sums into out_buf (natom
// rows of ms.stride).
int launch_moments_lmax(hp_part *p, int lmax, const MomShape &ms, double *part_buf, double *out_buf, bool molgrid) {
  if (ms.fast) return launch_moments(p, part_buf, out_buf, molgrid);
  DevFn f = p->fn();
  const int nq = ms.stride;
  const unsigned ngrp = (unsigned)((nq + HP_MOM_GROUP - 1) / HP_MOM_GROUP);
  const int nchunk = molgrid ? (int)((p->npoint + HP_MOM_CHUNK - 1) / HP_MOM_CHUNK) : 0;
  const dim3 g = molgrid ? dim3((unsigned)nchunk, (unsigned)p->natom, ngrp) : dim3((unsigned)p->nshell, 1, ngrp);
  const double *pm = p->promol;
#define HP_MOM_ANY(FK, MOL) \
  hipLaunchKernelGGL((k_moments_any<FK, MOL>), g, dim3(256), 0, p->stream, p->mol(), f, pm, lmax, nq, part_buf)
  if (molgrid) {
    if (p->fnkind == HP_FN_TABLE)
      HP_MOM_ANY(HP_FN_TABLE, true);
    else if (p->fnkind == HP_FN_SLATER)
      HP_MOM_ANY(HP_FN_SLATER, true);
    else if (p->fnkind == HP_FN_GAUSS)
      HP_MOM_ANY(HP_FN_GAUSS, true);
    else
      HP_MOM_ANY(HP_FN_GENERAL, true);
  } else {
    if (p->fnkind == HP_FN_TABLE)
      HP_MOM_ANY(HP_FN_TABLE, false);
    else if (p->fnkind == HP_FN_SLATER)
      HP_MOM_ANY(HP_FN_SLATER, false);
    else if (p->fnkind == HP_FN_GAUSS)
      HP_MOM_ANY(HP_FN_GAUSS, false);
    else
      HP_MOM_ANY(HP_FN_GENERAL, false);
  }
#undef HP_MOM_ANY
  HP_LAUNCHED();
  hipLaunchKernelGGL(k_moments_reduce_any, dim3(p->natom), dim3(128), 0, p->stream, p->mol(), molgrid ? 1 : 0, nchunk,
                     nq, part_buf, out_buf);
  HP_LAUNCHED();
  return 0;
}

int run_stage_gt2(hp_part *p, int precision, double thr, int min_stop, int max_iter) {
  HP_HIP(hipMemsetAsync(p->state, 0, sizeof(int), p->stream));
  while (true) {
    int rc = gt2_iteration(p, precision, thr, min_stop, max_iter);
    if (rc) return rc;
    int st[2];
    std::memcpy(st, p->h_out + 4, sizeof(st));
    p->niter = st[1];
    p->last_change = p->h_out[0];
    p->last_entropy = p->h_out[1];
    if ((int)p->h_out[2] != 0) return fail(HP_E_SOLVER, "a per-atom solver raised; see hp_part_status");
    if (st[0]) break;
  }
  return 0;
}

}  // namespace

// ---- static workspace of hp_part_moments_static ------------------------------
// Callers that cannot allocate device memory (code under stream capture,
// serving loops) get the moments' partial sums in this file-scope array
// instead of a buffer the library allocates.  Every GPU holds its own copy
// until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
#define HP_STATIC_WORKSPACE_DOUBLES (1 << 20)  // 8 MiB per GPU
static __device__ __align__(256) double scratch[HP_STATIC_WORKSPACE_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_workspace_mu;
double *g_workspace[kMaxDevices] = {};

// The device address of `scratch` on `device` (which the caller made current),
// looked up once per device.
int static_workspace(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return fail(HP_E_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_workspace_mu);
  if (!g_workspace[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(HP_E_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_workspace[device] = static_cast<double *>(ptr);
  }
  *out = g_workspace[device];
  return 0;
}

}  // namespace

extern "C" {

int hp_abi_version(void) { return HP_ABI_VERSION; }

const char *hp_last_error(void) { return g_err.c_str(); }

int hp_device_count(int *count) {
  if (!count) return fail(HP_E_ARG, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return fail(HP_E_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = n;
  return 0;
}

int hp_part_create(int device, const hp_molecule_desc *md, const hp_scheme_desc *sd, hp_part **out) {
  g_err.clear();
  if (!md || !sd || !out) return fail(HP_E_ARG, "NULL argument");
  *out = nullptr;
  const int A = md->natom, S = md->nshell;
  const int64_t N = md->npoint;
  if (A <= 0 || S <= 0 || N <= 0) return fail(HP_E_ARG, "natom, nshell and npoint must be positive");
  if (!md->coordinates || !md->pseudo_numbers || !md->points || !md->density || !md->mol_weights ||
      !md->atom_weights || !md->sph_weights || !md->shell_atom || !md->shell_offsets || !md->shell_radius ||
      !md->shell_rweight)
    return fail(HP_E_ARG, "molecule description has a NULL array");
  if (sd->scheme < 0 || sd->scheme > HP_SCHEME_HIRSHFELD) return fail(HP_E_ARG, "unknown scheme");
  if (sd->grid_type < 1 || sd->grid_type > 3) return fail(HP_E_ARG, "grid_type must be 1, 2 or 3");
  if (sd->grid_type >= 2 && sd->check_mono) return fail(HP_E_UNSUPP, "check_mono is defined on radial grids (grid_type 1)");
  const bool table = sd->scheme >= HP_SCHEME_ISA;
  if (table && sd->grid_type != 1 && sd->scheme != HP_SCHEME_HIRSHFELD)
    return fail(HP_E_UNSUPP, "ISA and Hirshfeld-I run on atom grids (grid_type 1)");
  if (!table && !sd->fn_offsets) return fail(HP_E_ARG, "fn_offsets is NULL");
  const bool mbis = sd->scheme == HP_SCHEME_MBIS;
  if (mbis && sd->solver != HP_SOLVER_MBIS) return fail(HP_E_ARG, "MBIS uses HP_SOLVER_MBIS");
  if (sd->scheme == HP_SCHEME_GISA && sd->solver != HP_SOLVER_QP) return fail(HP_E_ARG, "GISA uses HP_SOLVER_QP");
  const bool nlis = sd->scheme == HP_SCHEME_NLIS;
  if (nlis && sd->solver != HP_SOLVER_NLIS) return fail(HP_E_ARG, "NLIS uses HP_SOLVER_NLIS");
  if (nlis && !sd->fn_order) return fail(HP_E_ARG, "fn_order is NULL");
  if (nlis && sd->grid_type != 1) return fail(HP_E_UNSUPP, "NLIS/GMBIS on the device runs on atom grids (grid_type 1)");
  if (sd->scheme == HP_SCHEME_LISA &&
      !(sd->solver == HP_SOLVER_SC || sd->solver == HP_SOLVER_NEWTON || sd->solver == HP_SOLVER_MNEWTON))
    return fail(HP_E_UNSUPP, "LISA solver not available on the device");
  if (!table && sd->solver == HP_SOLVER_QP && !sd->qp_matrix) return fail(HP_E_ARG, "qp_matrix is NULL");
  if (!table && !mbis && !nlis && (!sd->fn_alpha || !sd->fn_order || !sd->fn_norm))
    return fail(HP_E_ARG, "basis arrays are NULL");
  if (sd->maxiter < 1 || sd->inner_maxiter < 1) return fail(HP_E_ARG, "iteration caps must be positive");
  if (sd->solver == HP_SOLVER_MNEWTON && sd->newton_linspace < 2) return fail(HP_E_ARG, "linspace size < 2");

  // validate grid structure: shells grouped by atom in order, offsets monotone
  std::vector<int> atom_shell_off(A + 1, 0);
  {
    int prev = -1;
    for (int s = 0; s < S; ++s) {
      const int a = md->shell_atom[s];
      if (a < 0 || a >= A || a < prev) return fail(HP_E_ARG, "shell_atom must be sorted atom indices");
      atom_shell_off[a + 1] += 1;
      prev = a;
    }
    for (int a = 0; a < A; ++a) {
      if (atom_shell_off[a + 1] == 0) return fail(HP_E_ARG, "every atom needs at least one radial shell");
      atom_shell_off[a + 1] += atom_shell_off[a];
    }
    if (md->shell_offsets[0] != 0 || md->shell_offsets[S] != N) return fail(HP_E_ARG, "shell_offsets must span the grid");
  }
  int maxL = 0, maxR = 0, maxK = 0;
  for (int s = 0; s < S; ++s) {
    const int64_t L = md->shell_offsets[s + 1] - md->shell_offsets[s];
    if (L <= 0) return fail(HP_E_ARG, "empty radial shell");
    maxL = std::max<int64_t>(maxL, L);
  }
  if (maxL > 1024) return fail(HP_E_UNSUPP, "more than 1024 points in one radial shell");
  for (int a = 0; a < A; ++a) maxR = std::max(maxR, atom_shell_off[a + 1] - atom_shell_off[a]);
  if (pick_ppl(maxR) == 0) return fail(HP_E_UNSUPP, "more than 512 radial points per atom");
  const int nfn = table ? 0 : sd->fn_offsets[A];
  if (!table && sd->fn_offsets[0] != 0) return fail(HP_E_ARG, "fn_offsets[0] must be 0");
  for (int a = 0; a < A && !table; ++a) {
    const int K = sd->fn_offsets[a + 1] - sd->fn_offsets[a];
    if (K <= 0) return fail(HP_E_ARG, "every atom needs at least one function");
    maxK = std::max(maxK, K);
  }
  // tabulated pro-atoms: knots per atom, records per atom and their values
  int ntab = 0;
  std::vector<int64_t> rec_val;
  if (table) {
    if (!sd->tab_offsets || !sd->tab_knots || !sd->tab_rweights) return fail(HP_E_ARG, "table arrays are NULL");
    if (sd->tab_offsets[0] != 0) return fail(HP_E_ARG, "tab_offsets[0] must be 0");
    for (int a = 0; a < A; ++a) {
      const int k0 = sd->tab_offsets[a], n = sd->tab_offsets[a + 1] - k0;
      const int nmin = sd->scheme == HP_SCHEME_ISA ? 4 : 2;
      if (n < nmin || n > HP_TAB_MAXN)
        return fail(n < nmin ? HP_E_ARG : HP_E_UNSUPP, "knots per atom out of range (ISA 4, otherwise 2, to 512)");
      for (int i = 1; i < n; ++i)
        if (!(sd->tab_knots[k0 + i] > sd->tab_knots[k0 + i - 1])) return fail(HP_E_ARG, "knots must ascend");
      if (sd->scheme == HP_SCHEME_ISA) {
        // the spherical averages of the atom's shells are the values at its knots
        const int s0 = atom_shell_off[a];
        if (n != atom_shell_off[a + 1] - s0) return fail(HP_E_ARG, "ISA knots must be the atom's radial shells");
        for (int i = 0; i < n; ++i)
          if (sd->tab_knots[k0 + i] != md->shell_radius[s0 + i])
            return fail(HP_E_ARG, "ISA knots must be the atom's radial shells");
      }
    }
    ntab = sd->tab_offsets[A];
    if (sd->scheme != HP_SCHEME_ISA) {
      if (!sd->rec_offsets || !sd->rec_charges || !sd->rec_rho || !sd->rec_coefs)
        return fail(HP_E_ARG, "record arrays are NULL");
      if (sd->rec_offsets[0] != 0) return fail(HP_E_ARG, "rec_offsets[0] must be 0");
      int64_t v = 0;
      for (int a = 0; a < A; ++a) {
        const int nr = sd->rec_offsets[a + 1] - sd->rec_offsets[a];
        if (nr < 1) return fail(HP_E_ARG, "every atom needs at least one record");
        if (sd->scheme == HP_SCHEME_HIRSHFELD && nr != 1) return fail(HP_E_ARG, "Hirshfeld takes one record per atom");
        const int n = sd->tab_offsets[a + 1] - sd->tab_offsets[a];
        for (int j = 0; j < nr; ++j) {
          rec_val.push_back(v);
          v += n;
        }
      }
    }
  }
  if (mbis && maxK > 7) return fail(HP_E_UNSUPP, "MBIS supports at most 7 shells per atom");
  if (nlis && maxK > HP_NLIS_KM) return fail(HP_E_UNSUPP, "NLIS supports at most 8 shells per atom");
  if (sd->grid_type >= 2 && maxK > HP_GT2_KM) return fail(HP_E_UNSUPP, "grid_type 2 supports at most 8 functions per atom");
  if (sd->grid_type >= 2 && sd->solver == HP_SOLVER_MNEWTON && sd->newton_linspace > HP_GT2_NLIN)
    return fail(HP_E_UNSUPP, "grid_type 2 line search supports at most 32 candidates");
  if (maxK > HP_KMAX) return fail(HP_E_UNSUPP, "too many functions per atom");
  int fnkind = table ? HP_FN_TABLE : HP_FN_SLATER;
  if (!mbis && !table) {
    bool all1 = true, all2 = true;
    for (int f = 0; f < nfn; ++f) {
      if (!(sd->fn_order[f] > 0.0) || !(nlis || sd->fn_alpha[f] >= 0.0))
        return fail(HP_E_ARG, "basis orders must be positive and exponents non-negative");
      all1 = all1 && sd->fn_order[f] == 1.0;
      all2 = all2 && sd->fn_order[f] == 2.0;
    }
    fnkind = all1 ? HP_FN_SLATER : (all2 ? HP_FN_GAUSS : HP_FN_GENERAL);
  }

  hp_part *p = new hp_part();
  p->device = device;
  p->natom = A;
  p->nshell = S;
  p->npoint = N;
  p->nfn = nfn;
  p->npar = mbis ? 2 * nfn : (nlis ? 3 * nfn : (table ? (sd->scheme == HP_SCHEME_ISA ? ntab : A) : nfn));
  p->ntab = ntab;
  p->scheme = sd->scheme;
  p->solver = sd->solver;
  // grid_type 3 runs the molecular-grid loop of grid_type 2 (core/base.py:490-499:
  // both set on_molgrid); it differs in where the moments integrate
  p->grid_type = sd->grid_type == 3 ? 2 : sd->grid_type;
  p->only_molgrid = sd->grid_type == 3;
  p->fnkind = fnkind;
  p->maxL = maxL;
  p->maxR = maxR;
  p->maxK = maxK;
  p->wshape = hp::weights_shape(maxL);
  p->threshold = sd->threshold;
  p->inner_threshold = sd->inner_threshold;
  p->dc = sd->density_cutoff;
  p->neg_cutoff = sd->negative_cutoff;
  p->pop_cutoff = sd->population_cutoff;
  p->rc = sd->cutoff_radius;
  p->tau = sd->newton_tau;
  p->nlin = sd->newton_linspace;
  p->check_mono = sd->check_mono;
  p->mixed_switch = sd->mixed_switch > 0.0 ? sd->mixed_switch : 0.0;
  p->mixed_refresh = sd->mixed_refresh > 0 ? sd->mixed_refresh : 0;
  p->tail_cull = sd->full_sum ? 0 : 1;
  p->maxiter = sd->maxiter;
  p->inner_maxiter = sd->inner_maxiter;
  p->h_shell_off.assign(md->shell_offsets, md->shell_offsets + S + 1);
  p->h_atom_shell_off = atom_shell_off;

  auto bail = [&](int rc) {
    free_part(p);
    return rc;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return bail(fail(HP_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
  e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(fail(HP_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  for (auto &ev : p->ev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) return bail(fail(HP_E_HIP, "hipEventCreate failed"));
  }
  for (auto &slot : p->slot_ev)
    for (auto &ev : slot) {
      e = hipEventCreate(&ev);
      if (e != hipSuccess) return bail(fail(HP_E_HIP, "hipEventCreate failed"));
    }
  e = hipHostMalloc((void **)&p->h_out, 8 * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) return bail(fail(HP_E_HIP, "hipHostMalloc failed"));
  hipStream_t s = p->stream;

  // SoA copies of coordinates and points
  std::vector<double> ax(A), ay(A), az(A);
  for (int a = 0; a < A; ++a) {
    ax[a] = md->coordinates[3 * a];
    ay[a] = md->coordinates[3 * a + 1];
    az[a] = md->coordinates[3 * a + 2];
  }
  std::vector<double> gx(N), gy(N), gz(N);
  for (int64_t i = 0; i < N; ++i) {
    gx[i] = md->points[3 * i];
    gy[i] = md->points[3 * i + 1];
    gz[i] = md->points[3 * i + 2];
  }
  std::vector<int> shell_atom(md->shell_atom, md->shell_atom + S);
  int rc = 0;
  if ((rc = dupload(&p->cx, ax.data(), A, s)) || (rc = dupload(&p->cy, ay.data(), A, s)) ||
      (rc = dupload(&p->cz, az.data(), A, s)) || (rc = dupload(&p->px, gx.data(), N, s)) ||
      (rc = dupload(&p->py, gy.data(), N, s)) || (rc = dupload(&p->pz, gz.data(), N, s)) ||
      (rc = dupload(&p->rho, md->density, N, s)) || (rc = dupload(&p->wmol, md->mol_weights, N, s)) ||
      (rc = dupload(&p->wat, md->atom_weights, N, s)) || (rc = dupload(&p->wsph, md->sph_weights, N, s)) ||
      (rc = dupload(&p->shell_r, md->shell_radius, S, s)) || (rc = dupload(&p->shell_w, md->shell_rweight, S, s)) ||
      (rc = dupload(&p->pseudo, md->pseudo_numbers, A, s)) || (rc = dupload(&p->shell_atom, shell_atom.data(), S, s)) ||
      (rc = dupload(&p->shell_off, md->shell_offsets, S + 1, s)) ||
      (rc = dupload(&p->atom_shell_off, atom_shell_off.data(), A + 1, s)))
    return bail(rc);
  std::vector<int> fn_off(A + 1, 0);
  if (!table) fn_off.assign(sd->fn_offsets, sd->fn_offsets + A + 1);
  if ((rc = dupload(&p->fn_off, fn_off.data(), A + 1, s))) return bail(rc);
  // host copies live until the stream synchronisation at the end of create
  std::vector<int> toff, guide, roff, rch;
  std::vector<float2> gparam;
  if (table) {
    toff.assign(sd->tab_offsets, sd->tab_offsets + A + 1);
    // interval guide per atom: bin j of log2 r starts at the interval that
    // holds rmin 2^(j / bins-per-unit) (the kernels walk to the exact one)
    guide.resize((size_t)A * HP_TAB_GUIDE);
    gparam.resize(A);
    for (int a = 0; a < A; ++a) {
      const double *x = sd->tab_knots + toff[a];
      const int n = toff[a + 1] - toff[a];
      const double u0 = std::log2(x[0]), u1 = std::log2(x[n - 1]);
      const double per = (u1 > u0) ? (HP_TAB_GUIDE - 1) / (u1 - u0) : 0.0;
      gparam[a] = make_float2((float)u0, (float)per);
      for (int j = 0; j < HP_TAB_GUIDE; ++j) {
        const double r = per > 0.0 ? std::exp2(u0 + j / per) : x[0];
        guide[(size_t)a * HP_TAB_GUIDE + j] = table_interval(x, n, r);
      }
    }
    if ((rc = dupload(&p->tab_guide, guide.data(), guide.size(), s)) ||
        (rc = dupload(&p->tab_gparam, gparam.data(), gparam.size(), s)))
      return bail(rc);
    if ((rc = dupload(&p->tab_off, toff.data(), A + 1, s)) || (rc = dupload(&p->tab_x, sd->tab_knots, ntab, s)) ||
        (rc = dupload(&p->tab_w, sd->tab_rweights, ntab, s)) || (rc = dalloc(&p->tab_c, (size_t)4 * ntab)) ||
        (rc = dalloc(&p->cur_rho, ntab)) || (rc = dalloc(&p->pop_part, S)))
      return bail(rc);
    if (sd->scheme != HP_SCHEME_ISA) {
      const int nrec = sd->rec_offsets[A];
      const int64_t nval = rec_val.empty() ? 0 : rec_val.back() + (sd->tab_offsets[A] - sd->tab_offsets[A - 1]);
      roff.assign(sd->rec_offsets, sd->rec_offsets + A + 1);
      rch.assign(sd->rec_charges, sd->rec_charges + nrec);
      if ((rc = dupload(&p->rec_off, roff.data(), A + 1, s)) || (rc = dupload(&p->rec_charge, rch.data(), nrec, s)) ||
          (rc = dupload(&p->rec_val, rec_val.data(), rec_val.size(), s)) ||
          (rc = dupload(&p->rec_rho, sd->rec_rho, nval, s)) || (rc = dupload(&p->rec_c, sd->rec_coefs, 4 * nval, s)))
        return bail(rc);
      if (sd->scheme == HP_SCHEME_HIRSHFELD) {
        // the fixed pro-atoms: record a is atom a's table
        const hipError_t ce =
            hipMemcpyAsync(p->tab_c, p->rec_c, sizeof(double) * 4 * (size_t)ntab, hipMemcpyDeviceToDevice, s);
        if (ce != hipSuccess) return bail(fail(HP_E_HIP, std::string("table copy: ") + hipGetErrorString(ce)));
      }
    }
  }
  const std::vector<int> grp_first = hp::weights_groups(atom_shell_off, p->wshape.g);
  p->ngroup = (int)grp_first.size();
  if ((rc = dupload(&p->grp_first, grp_first.data(), grp_first.size(), s))) return bail(rc);
  if (nlis) {
    // n / (4 pi Gamma(3/n)) per function (n fixed: the solver updates N and S only)
    std::vector<double> g(nfn);
    for (int f = 0; f < nfn; ++f) g[f] = sd->fn_order[f] / (4.0 * M_PI * std::tgamma(3.0 / sd->fn_order[f]));
    p->h_order.assign(sd->fn_order, sd->fn_order + nfn);
    if ((rc = dupload(&p->fn_norm, g.data(), nfn, s)) || (rc = dupload(&p->fn_order, sd->fn_order, nfn, s)))
      return bail(rc);
  } else if (!mbis) {
    if ((rc = dupload(&p->fn_norm, sd->fn_norm, nfn, s)) || (rc = dupload(&p->fn_alpha, sd->fn_alpha, nfn, s)) ||
        (rc = dupload(&p->fn_order, sd->fn_order, nfn, s)))
      return bail(rc);
  } else {
    std::vector<double> ones(nfn, 1.0);
    if ((rc = dupload(&p->fn_order, ones.data(), nfn, s))) return bail(rc);
  }
  if (sd->solver == HP_SOLVER_QP) {
    std::vector<int64_t> qoff(A + 1, 0);
    for (int a = 0; a < A; ++a) {
      const int64_t K = fn_off[a + 1] - fn_off[a];
      qoff[a + 1] = qoff[a] + K * K;
    }
    if ((rc = dupload(&p->qpP, sd->qp_matrix, qoff[A], s)) || (rc = dupload(&p->qp_off, qoff.data(), A + 1, s)))
      return bail(rc);
  }
  // neighbour lists for the cut-off: atoms sorted by distance to each atom,
  // and per shell the number of atoms that can reach any of its points
  std::vector<int> nb_h, shell_nb_h;
  if (p->rc > 0.0) {
    std::vector<int> &nb = nb_h;
    nb.resize((size_t)A * A);
    std::vector<double> dist((size_t)A * A);
    for (int a = 0; a < A; ++a) {
      std::vector<std::pair<double, int>> v(A);
      for (int b = 0; b < A; ++b) {
        const double dx = ax[a] - ax[b], dy = ay[a] - ay[b], dz = az[a] - az[b];
        v[b] = {std::sqrt(dx * dx + dy * dy + dz * dz), b};
      }
      std::sort(v.begin(), v.end());
      for (int t = 0; t < A; ++t) {
        nb[(size_t)a * A + t] = v[t].second;
        dist[(size_t)a * A + t] = v[t].first;
      }
    }
    std::vector<int> &shell_nb = shell_nb_h;
    shell_nb.resize(S);
    for (int si = 0; si < S; ++si) {
      const int a = shell_atom[si];
      double rmax = 0.0;
      for (int64_t i = md->shell_offsets[si]; i < md->shell_offsets[si + 1]; ++i) {
        const double dx = gx[i] - ax[a], dy = gy[i] - ay[a], dz = gz[i] - az[a];
        rmax = std::max(rmax, std::sqrt(dx * dx + dy * dy + dz * dz));
      }
      const double reach = rmax + p->rc + 1e-9 * (1.0 + rmax + p->rc);
      const double *d = &dist[(size_t)a * A];
      shell_nb[si] = (int)(std::upper_bound(d, d + A, reach) - d);
    }
    if ((rc = dupload(&p->nb_list, nb.data(), nb.size(), s)) || (rc = dupload(&p->shell_nb, shell_nb.data(), S, s)))
      return bail(rc);
  }
  // algorithmic work of one weights pass: function evaluations per point
  // over the atoms that reach its shell
  if (p->rc > 0.0) {
    double ev = 0.0;
    for (int si = 0; si < S; ++si) {
      const int a = shell_atom[si];
      const double L = (double)(md->shell_offsets[si + 1] - md->shell_offsets[si]);
      int64_t kf = 0;
      for (int t = 0; t < shell_nb_h[si]; ++t) {
        const int b = nb_h[(size_t)a * A + t];
        kf += fn_off[b + 1] - fn_off[b];
      }
      ev += L * (double)kf;
    }
    p->evals_per_pass = ev;
  } else {
    // tabulated pro-atoms: one spline evaluation per point and atom
    p->evals_per_pass = (double)N * (double)(table ? A : nfn);
  }
  // state
  p->hist_cap = 2 * p->maxiter + 64;
  if ((rc = dalloc(&p->state, 2))) return bail(rc);
  if ((rc = dalloc(&p->coef, nfn)) || (rc = dalloc(&p->alpha, nfn)) || (rc = dalloc(&p->par, p->npar)) ||
      (rc = dalloc(&p->old_par, p->npar)) || (rc = dalloc(&p->rhobar, S)) || (rc = dalloc(&p->ent_part, S)) ||
      (rc = dalloc(&p->charges, A)) || (rc = dalloc(&p->change_part, A)) || (rc = dalloc(&p->out, 4)) ||
      (rc = dalloc(&p->inner_niter, A)) || (rc = dalloc(&p->status, A)) || (rc = dalloc(&p->status_all, A)) ||
      (rc = dalloc(&p->hpar, (size_t)p->hist_cap * p->npar)) || (rc = dalloc(&p->hch, (size_t)p->hist_cap * A)) ||
      (rc = dalloc(&p->hent, p->hist_cap)) || (rc = dalloc(&p->hchg, p->hist_cap)) || (rc = dalloc(&p->corr, S)) ||
      (rc = dalloc(&p->evals_done, HP_EVAL_SLOTS)))
    return bail(rc);
  if (p->grid_type == 2 && table) {
    // Hirshfeld on the molecular grid: the promolecule for the populations
    if ((rc = dalloc(&p->promol, N))) return bail(rc);
  } else if (p->grid_type == 2) {
    p->gt2_nchunk = (int)((N + HP_GT2_CHUNK - 1) / HP_GT2_CHUNK);
    if ((rc = dalloc(&p->promol, N)) || (rc = dalloc(&p->gt2, A)) ||
        (rc = dalloc(&p->gt2_partial, (size_t)A * p->gt2_nchunk * HP_GT2_NV)) || (rc = dalloc(&p->gt2_active, 1)) ||
        (rc = dalloc(&p->gt2_nhess, A)))
      return bail(rc);
  }
  e = hipMemsetAsync(p->state, 0, sizeof(int) * 2, s);
  if (e == hipSuccess) e = hipMemsetAsync(p->status_all, 0, sizeof(int) * A, s);
  if (e == hipSuccess) e = hipMemsetAsync(p->status, 0, sizeof(int) * A, s);
  if (e == hipSuccess) e = hipMemsetAsync(p->inner_niter, 0, sizeof(int) * A, s);
  if (e == hipSuccess) e = hipMemsetAsync(p->charges, 0, sizeof(double) * A, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return bail(fail(HP_E_HIP, std::string("initialisation: ") + hipGetErrorString(e)));
  *out = p;
  return 0;
}

int hp_part_destroy(hp_part *part) {
  if (!part) return fail(HP_E_ARG, "NULL part");
  (void)hipSetDevice(part->device);
  (void)hipStreamSynchronize(part->stream);
  free_part(part);
  return 0;
}

int hp_part_set_propars(hp_part *p, const double *propars, int64_t n) {
  if (!p || !propars) return fail(HP_E_ARG, "NULL argument");
  if (n != p->npar) return fail(HP_E_ARG, "propars length mismatch");
  if (p->scheme == HP_SCHEME_NLIS)
    for (int f = 0; f < p->nfn; ++f)
      if (propars[3 * f + 2] != p->h_order[f]) return fail(HP_E_ARG, "NLIS powers n differ from the scheme's fn_order");
  bool nonneg = true;
  for (int64_t k = 0; k < n; ++k) nonneg = nonneg && propars[k] >= 0.0;
  p->params_nonneg = nonneg ? 1 : 0;
  HP_HIP(hipSetDevice(p->device));
  HP_HIP(hipMemcpyAsync(p->par, propars, sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
  HP_HIP(hipMemcpyAsync(p->old_par, propars, sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  p->niter = 0;
  p->weights_ready = 0;
  return 0;
}

int hp_part_weights(hp_part *p, double *rhobar, double *entropy) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  if (rhobar && p->grid_type != 1) return fail(HP_E_UNSUPP, "spherical averages exist for grid_type 1 only");
  HP_HIP(hipSetDevice(p->device));
  hipStream_t s = p->stream;
  // the kernels skip their work while the device stop flag is set
  HP_HIP(hipMemsetAsync(p->state, 0, sizeof(int), s));
  if (p->tabulated()) {
    HP_HIP(hipMemsetAsync(p->status, 0, sizeof(int) * p->natom, s));
    launch_table_begin(p);
    HP_LAUNCHED();
    HP_HIP(hipEventRecord(p->ev[0], s));
    launch_table_weights(p);
  } else {
    const int n = std::max(p->npar, p->nfn);
    hipLaunchKernelGGL(k_iter_begin, dim3((n + 255) / 256), dim3(256), 0, s, p->state, table_kind(p),
                       p->npar, p->nfn, p->par, p->old_par, p->fn_norm, p->fn_alpha, p->coef, p->alpha);
    HP_LAUNCHED();
    HP_HIP(hipEventRecord(p->ev[0], s));
    launch_weights<double>(p, p->grid_type == 2);
  }
  HP_LAUNCHED();
  HP_HIP(hipEventRecord(p->ev[1], s));
  hipLaunchKernelGGL(k_sum_shells, dim3(1), dim3(256), 0, s, p->nshell, p->ent_part, p->out + 1);
  HP_LAUNCHED();
  if (rhobar) HP_HIP(hipMemcpyAsync(rhobar, p->rhobar, sizeof(double) * p->nshell, hipMemcpyDeviceToHost, s));
  HP_HIP(hipMemcpyAsync(p->h_out + 1, p->out + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  std::vector<int> st(p->tabulated() ? p->natom : 0);
  if (!st.empty()) HP_HIP(hipMemcpyAsync(st.data(), p->status, sizeof(int) * st.size(), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  for (int v : st)
    if (v & HP_ST_NORECORD) return fail(HP_E_SOLVER, "no pro-atom for an atom's charge; see hp_part_status");
  float ms = 0.f;
  HP_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
  p->k_ms[HP_KERNEL_WEIGHTS] += ms;
  p->k_n[HP_KERNEL_WEIGHTS] += 1;
  p->last_entropy = p->h_out[1];
  if (entropy) *entropy = p->h_out[1];
  p->weights_ready = 1;
  return 0;
}

int hp_part_run(hp_part *p, int precision, int n_polish, int *niter) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  if (precision != HP_PREC_FP64 && precision != HP_PREC_MIXED) return fail(HP_E_ARG, "unknown precision");
  // the histories hold 2*maxiter + 64 iterations
  if (precision == HP_PREC_MIXED && (n_polish < 0 || n_polish > p->maxiter + 64))
    return fail(HP_E_ARG, "n_polish must be in [0, maxiter + 64]");
  if (p->scheme == HP_SCHEME_HIRSHFELD) return fail(HP_E_UNSUPP, "Hirshfeld does not iterate: use hp_part_weights");
  if (p->tabulated() && precision != HP_PREC_FP64) return fail(HP_E_UNSUPP, "ISA and Hirshfeld-I run in FP64");
  HP_HIP(hipSetDevice(p->device));
  for (int k = 0; k < 3; ++k) {
    p->k_ms[k] = 0.0;
    p->k_n[k] = 0;
  }
  HP_HIP(hipMemsetAsync(p->status_all, 0, sizeof(int) * p->natom, p->stream));
  HP_HIP(hipMemsetAsync(p->state, 0, 2 * sizeof(int), p->stream));
  if (p->gt2_nhess) HP_HIP(hipMemsetAsync(p->gt2_nhess, 0, sizeof(int) * p->natom, p->stream));
  HP_HIP(hipMemsetAsync(p->evals_done, 0, HP_EVAL_SLOTS * sizeof(unsigned long long), p->stream));
  p->niter = 0;
  int rc = 0;
  auto stage = [&](int prec, double thr, int min_stop, int max_iter) {
    return p->grid_type == 2 ? run_stage_gt2(p, prec, thr, min_stop, max_iter)
                             : run_stage(p, prec, thr, min_stop, max_iter);
  };
  if (precision == HP_PREC_FP64) {
    // iterstock.py:172-188: stop when change < threshold or counter >= maxiter
    rc = stage(HP_PREC_FP64, p->threshold, 1, p->maxiter);
  } else {
    if (p->grid_type == 2) {
      // FP32 stage until change < max(threshold, switch level (1e-5 unless
      // set)), then FP64 iterations (at least n_polish) until change < threshold
      const double sw = p->mixed_switch > 0.0 ? p->mixed_switch : 1e-5;
      rc = run_stage_gt2(p, HP_PREC_MIXED, std::max(p->threshold, sw), 1, p->maxiter);
    } else {
      // FP32 stage (with the FP64 defect correction every mixed_refresh
      // iterations) until change < max(threshold * HP_MIXED_MARGIN, switch
      // level); the iteration that passes is discarded and redone in FP64,
      // and FP64 iterations (at least n_polish) follow until change < threshold
      const double thr1 = std::max(p->threshold * HP_MIXED_MARGIN, p->mixed_switch);
      rc = run_stage(p, HP_PREC_MIXED, thr1, 1, p->maxiter, 1);
    }
    // maxiter caps the total unless the polish needs more
    if (!rc) {
      const int start = p->niter;
      const int np = std::max(1, n_polish);
      rc = stage(HP_PREC_FP64, p->threshold, start + np, std::max(p->maxiter, start + np));
    }
  }
  if (niter) *niter = p->niter;
  return rc;
}

int hp_part_results(hp_part *p, double *propars, double *charges, double *change, double *entropy) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  HP_HIP(hipSetDevice(p->device));
  if (propars) HP_HIP(hipMemcpyAsync(propars, p->par, sizeof(double) * p->npar, hipMemcpyDeviceToHost, p->stream));
  if (charges) HP_HIP(hipMemcpyAsync(charges, p->charges, sizeof(double) * p->natom, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  if (change) *change = p->last_change;
  if (entropy) *entropy = p->last_entropy;
  return 0;
}

int hp_part_history(hp_part *p, double *hpropars, double *hcharges, double *hentropies, double *hchanges) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  HP_HIP(hipSetDevice(p->device));
  const size_t n = (size_t)p->niter;
  if (hpropars && n)
    HP_HIP(hipMemcpyAsync(hpropars, p->hpar, sizeof(double) * n * p->npar, hipMemcpyDeviceToHost, p->stream));
  if (hcharges && n)
    HP_HIP(hipMemcpyAsync(hcharges, p->hch, sizeof(double) * n * p->natom, hipMemcpyDeviceToHost, p->stream));
  if (hentropies && n) HP_HIP(hipMemcpyAsync(hentropies, p->hent, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
  if (hchanges && n) HP_HIP(hipMemcpyAsync(hchanges, p->hchg, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  return 0;
}

int hp_part_status(hp_part *p, int32_t *inner_niter, int32_t *status_last, int32_t *status_all) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  HP_HIP(hipSetDevice(p->device));
  if (inner_niter)
    HP_HIP(hipMemcpyAsync(inner_niter, p->inner_niter, sizeof(int) * p->natom, hipMemcpyDeviceToHost, p->stream));
  if (status_last)
    HP_HIP(hipMemcpyAsync(status_last, p->status, sizeof(int) * p->natom, hipMemcpyDeviceToHost, p->stream));
  if (status_all)
    HP_HIP(hipMemcpyAsync(status_all, p->status_all, sizeof(int) * p->natom, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  return 0;
}

int hp_part_spherical_averages(hp_part *p, double *rhobar) {
  if (!p || !rhobar) return fail(HP_E_ARG, "NULL argument");
  if (p->grid_type != 1) return fail(HP_E_UNSUPP, "spherical averages exist for grid_type 1 only");
  HP_HIP(hipSetDevice(p->device));
  HP_HIP(hipMemcpyAsync(rhobar, p->rhobar, sizeof(double) * p->nshell, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  return 0;
}

static int point_weights(hp_part *p, int atom, int64_t p0, int64_t n, double *out) {
  HP_HIP(hipSetDevice(p->device));
  if (!p->scratch_pts) {
    int rc = dalloc(&p->scratch_pts, (size_t)p->npoint);
    if (rc) return rc;
  }
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  DevFn f = p->fn();
  if (p->fnkind == HP_FN_TABLE)
    hipLaunchKernelGGL((k_point_weights<HP_FN_TABLE>), grid, block, 0, p->stream, p->mol(), f, atom, p0, n,
                       p->scratch_pts);
  else if (p->fnkind == HP_FN_SLATER)
    hipLaunchKernelGGL((k_point_weights<HP_FN_SLATER>), grid, block, 0, p->stream, p->mol(), f, atom, p0, n,
                       p->scratch_pts);
  else if (p->fnkind == HP_FN_GAUSS)
    hipLaunchKernelGGL((k_point_weights<HP_FN_GAUSS>), grid, block, 0, p->stream, p->mol(), f, atom, p0, n,
                       p->scratch_pts);
  else
    hipLaunchKernelGGL((k_point_weights<HP_FN_GENERAL>), grid, block, 0, p->stream, p->mol(), f, atom, p0, n,
                       p->scratch_pts);
  HP_LAUNCHED();
  HP_HIP(hipMemcpyAsync(out, p->scratch_pts, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  return 0;
}

int hp_part_at_weights(hp_part *p, int atom, double *out, int64_t n) {
  if (!p || !out) return fail(HP_E_ARG, "NULL argument");
  if (atom < 0 || atom >= p->natom) return fail(HP_E_ARG, "atom index out of range");
  if (p->niter == 0 && !p->weights_ready) return fail(HP_E_ARG, "no weights yet: run the partitioning first");
  int64_t p0 = p->h_shell_off[p->h_atom_shell_off[atom]];
  int64_t p1 = p->h_shell_off[p->h_atom_shell_off[atom + 1]];
  if (p->grid_type == 2) {  // weights on the whole molecular grid
    p0 = 0;
    p1 = p->npoint;
  }
  if (n != p1 - p0) return fail(HP_E_ARG, "output length must equal the weight-array size");
  return point_weights(p, atom, p0, n, out);
}

int hp_part_promoldens(hp_part *p, double *out, int64_t n) {
  if (!p || !out) return fail(HP_E_ARG, "NULL argument");
  if (n != p->npoint) return fail(HP_E_ARG, "output length must equal npoint");
  if (p->niter == 0 && !p->weights_ready) return fail(HP_E_ARG, "no promolecule yet: run the partitioning first");
  return point_weights(p, -1, 0, n, out);
}

int hp_part_moments(hp_part *p, int lmax, double *cart, double *pure, double *radial) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  if (lmax < 0) return fail(HP_E_ARG, "lmax must be >= 0");
  if (p->niter == 0 && !p->weights_ready) return fail(HP_E_ARG, "run the partitioning first");
  HP_HIP(hipSetDevice(p->device));
  int rc = 0;
  const MomShape ms = moment_shape(lmax);
  const size_t npart = moments_partials(p, p->only_molgrid) * (size_t)ms.stride;
  const size_t nout = (size_t)p->natom * (size_t)ms.stride;
  if (p->mom_part_cap < npart) {
    if (p->mom_part) HP_HIP(hipFree(p->mom_part));
    p->mom_part = nullptr;
    p->mom_part_cap = 0;
    if ((rc = dalloc(&p->mom_part, npart))) return rc;
    p->mom_part_cap = npart;
  }
  if (p->mom_out_cap < nout) {
    if (p->mom_out) HP_HIP(hipFree(p->mom_out));
    p->mom_out = nullptr;
    p->mom_out_cap = 0;
    if ((rc = dalloc(&p->mom_out, nout))) return rc;
    p->mom_out_cap = nout;
  }
  if ((rc = launch_moments_lmax(p, lmax, ms, p->mom_part, p->mom_out, p->only_molgrid))) return rc;
  std::vector<double> h(nout);
  std::vector<double> z(p->natom);
  HP_HIP(hipMemcpyAsync(h.data(), p->mom_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipMemcpyAsync(z.data(), p->pseudo, sizeof(double) * p->natom, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  for (int a = 0; a < p->natom; ++a) {
    const double *v = &h[(size_t)a * ms.stride];
    if (cart) {
      double *c = cart + (size_t)a * ms.ncart;
      for (int i = 0; i < ms.ncart; ++i) c[i] = -v[ms.src_cart + i];
      c[0] += z[a];
    }
    if (pure) {
      double *u = pure + (size_t)a * ms.npure;
      for (int i = 0; i < ms.npure; ++i) u[i] = -v[ms.src_pure + i];
      u[0] += z[a];
    }
    if (radial) {
      double *r = radial + (size_t)a * ms.nrad;
      for (int i = 0; i < ms.nrad; ++i) r[i] = v[ms.src_rad + i];
    }
  }
  return 0;
}

int hp_part_moments_static(hp_part *p, int lmax, double *cart, double *pure, double *radial) {
  if (!p) return fail(HP_E_ARG, "NULL part");
  if (lmax < 0) return fail(HP_E_ARG, "lmax must be >= 0");
  if (p->niter == 0 && !p->weights_ready) return fail(HP_E_ARG, "run the partitioning first");
  // partials, per-atom sums, then cart | pure | radial of every atom
  const MomShape ms = moment_shape(lmax);
  const size_t A = (size_t)p->natom;
  const size_t npart = moments_partials(p, p->only_molgrid) * (size_t)ms.stride;
  const size_t nlay = A * (size_t)(ms.ncart + ms.npure + ms.nrad);
  const size_t need = npart + A * (size_t)ms.stride + nlay;
  if (need > (size_t)HP_STATIC_WORKSPACE_DOUBLES)
    return fail(HP_E_ARG, "moments of this molecule need " + std::to_string(need) + " doubles of workspace, beyond the " +
                              std::to_string(HP_STATIC_WORKSPACE_DOUBLES) +
                              " of the static workspace (partial rows and natom rows of " +
                              std::to_string(ms.stride) + " sums, and the natom x " +
                              std::to_string(ms.ncart + ms.npure + ms.nrad) + " output); use hp_part_moments");
  HP_HIP(hipSetDevice(p->device));
  double *ws = nullptr;
  if (int rc = static_workspace(p->device, &ws)) return rc;
  double *part_buf = ws, *out_buf = ws + npart, *lay = out_buf + A * (size_t)ms.stride;
  if (int rc = launch_moments_lmax(p, lmax, ms, part_buf, out_buf, p->only_molgrid)) return rc;
  hipLaunchKernelGGL(k_moments_layout, dim3((p->natom + 63) / 64), dim3(64), 0, p->stream, p->natom, ms.stride,
                     ms.src_cart, ms.ncart, ms.src_pure, ms.npure, ms.src_rad, ms.nrad, out_buf, p->pseudo, lay);
  HP_LAUNCHED();
  const double *lay_pure = lay + A * ms.ncart, *lay_rad = lay_pure + A * ms.npure;
  if (cart) HP_HIP(hipMemcpyAsync(cart, lay, sizeof(double) * A * ms.ncart, hipMemcpyDeviceToHost, p->stream));
  if (pure) HP_HIP(hipMemcpyAsync(pure, lay_pure, sizeof(double) * A * ms.npure, hipMemcpyDeviceToHost, p->stream));
  if (radial) HP_HIP(hipMemcpyAsync(radial, lay_rad, sizeof(double) * A * ms.nrad, hipMemcpyDeviceToHost, p->stream));
  HP_HIP(hipStreamSynchronize(p->stream));
  return 0;
}

static int glisa_setup(hp_part *p) {
  if (p->gl_G) return 0;
  const int M = p->nfn, Mp = ((M + 15) / 16) * 16, T = Mp / 16;
  p->gl_M = M;
  p->gl_Mp = Mp;
  p->gl_npair = T * (T + 1) / 2;
  p->gl_nchunk = (int)((p->npoint + HP_GLISA_CHUNK - 1) / HP_GLISA_CHUNK);
  p->gl_hnchunk = (int)((p->npoint + HP_GLISA_HCHUNK - 1) / HP_GLISA_HCHUNK);
  p->gl_nblock = (int)((p->npoint + 255) / 256);
  std::vector<int2> pairs;
  for (int I = 0; I < T; ++I)
    for (int J = I; J < T; ++J) pairs.push_back(make_int2(I, J));
  int rc = 0;
  if ((rc = dalloc(&p->gl_G, (size_t)p->npoint * Mp)) || (rc = dalloc(&p->gl_D, (size_t)p->npoint)) ||
      (rc = dalloc(&p->gl_E, (size_t)p->npoint)) || (rc = dalloc(&p->gl_obj, 2 * (size_t)p->gl_nblock)) ||
      (rc = dalloc(&p->gl_grad, (size_t)std::max(p->gl_nchunk, p->gl_hnchunk) * M)) ||
      (rc = dalloc(&p->gl_hess, (size_t)p->gl_npair * p->gl_hnchunk * 256)) ||
      (rc = dalloc(&p->gl_out, 2 + (size_t)M + (size_t)M * M)) || (rc = dalloc(&p->gl_coef, 2 * (size_t)M)) ||
      (rc = dalloc(&p->gl_pro, 2 * (size_t)p->natom * (p->gl_nchunk + 1))) ||
      (rc = dupload(&p->gl_pairs, pairs.data(), pairs.size(), p->stream)))
    return rc;
  return 0;
}

int hp_part_global_terms(hp_part *p, const double *coefs, int nderiv, double *objective, double *gradient,
                         double *hessian, double *min_promol) {
  if (!p || !coefs) return fail(HP_E_ARG, "NULL argument");
  if (p->scheme != HP_SCHEME_LISA) return fail(HP_E_UNSUPP, "the global terms need a linear-coefficient basis (LISA)");
  if (nderiv < 0 || nderiv > 2) return fail(HP_E_ARG, "nderiv must be 0, 1 or 2");
  HP_HIP(hipSetDevice(p->device));
  if (int rc = glisa_setup(p)) return rc;
  hipStream_t s = p->stream;
  const int M = p->gl_M, Mp = p->gl_Mp;
  HP_HIP(hipMemcpyAsync(p->gl_coef, coefs, sizeof(double) * M, hipMemcpyHostToDevice, s));
  DevFn f = p->fn();
  f.alpha = p->fn_alpha;
  const dim3 g1((unsigned)p->gl_nblock), b1(256);
#define HP_GL_POINTS(FK)                                                                                          \
  hipLaunchKernelGGL((k_glisa_points<FK>), g1, b1, 0, s, p->mol(), f, p->fn_norm, p->gl_coef, M, Mp, p->dc, p->gl_G, \
                     p->gl_D, p->gl_E, p->gl_obj)
  if (p->fnkind == HP_FN_SLATER)
    HP_GL_POINTS(HP_FN_SLATER);
  else if (p->fnkind == HP_FN_GAUSS)
    HP_GL_POINTS(HP_FN_GAUSS);
  else
    HP_GL_POINTS(HP_FN_GENERAL);
#undef HP_GL_POINTS
  HP_LAUNCHED();
  // with the Hessian of the LDS kernel the gradient comes from its staged
  // rows (partials per Hessian chunk), else from its own pass over G
  const bool fused_grad = nderiv == 2 && Mp <= 112;
  const int gchunk = fused_grad ? p->gl_hnchunk : p->gl_nchunk;
  if (nderiv >= 1 && !fused_grad) {
    hipLaunchKernelGGL(k_glisa_grad, dim3((unsigned)p->gl_nchunk), dim3(256), 0, s, p->npoint, M, Mp, p->gl_G, p->gl_E,
                       p->gl_grad);
    HP_LAUNCHED();
  }
  if (nderiv == 2) {
    HP_HIP(hipEventRecord(p->ev[0], s));
    // every tile pair of a chunk in one workgroup while the pairs fit its
    // wavefronts' registers (Mp <= 112: at most 28 pairs, 7 per wavefront),
    // else one wavefront per pair and chunk
    if (Mp <= 112)
      hipLaunchKernelGGL((k_glisa_hess_lds<16, 112, 7>), dim3((unsigned)p->gl_hnchunk), dim3(256), 0, s, p->npoint,
                         Mp, p->gl_npair, p->gl_pairs, p->gl_G, p->gl_D, p->gl_hess, M, p->gl_E, p->gl_grad);
    else
      hipLaunchKernelGGL(k_glisa_hess, dim3((unsigned)p->gl_npair, (unsigned)p->gl_hnchunk), dim3(64), 0, s,
                         p->npoint, Mp, p->gl_pairs, p->gl_G, p->gl_D, p->gl_hess);
    HP_LAUNCHED();
    HP_HIP(hipEventRecord(p->ev[1], s));
  }
  const int nt = std::max({1, nderiv >= 1 ? M : 0, nderiv == 2 ? p->gl_npair * 256 : 0});
  hipLaunchKernelGGL(k_glisa_reduce, dim3((unsigned)(1 + (nt + 255) / 256)), dim3(256), 0, s, p->gl_nblock, p->gl_obj,
                     gchunk, M, nderiv >= 1 ? 1 : 0, p->gl_grad, nderiv == 2 ? p->gl_npair : 0, p->gl_pairs,
                     p->gl_hnchunk, nderiv == 2 ? p->gl_hess : nullptr, p->gl_out);
  HP_LAUNCHED();
  if (objective) HP_HIP(hipMemcpyAsync(objective, p->gl_out, sizeof(double), hipMemcpyDeviceToHost, s));
  if (min_promol)
    HP_HIP(hipMemcpyAsync(min_promol, p->gl_out + 1 + M + (size_t)M * M, sizeof(double), hipMemcpyDeviceToHost, s));
  if (gradient && nderiv >= 1)
    HP_HIP(hipMemcpyAsync(gradient, p->gl_out + 1, sizeof(double) * M, hipMemcpyDeviceToHost, s));
  if (hessian && nderiv == 2)
    HP_HIP(hipMemcpyAsync(hessian, p->gl_out + 1 + M, sizeof(double) * M * M, hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  if (nderiv == 2) {
    float ms = 0.f;
    HP_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
    p->k_ms[HP_KERNEL_GLISA_HESS] += ms;
    p->k_n[HP_KERNEL_GLISA_HESS] += 1;
  }
  return 0;
}

int hp_part_global_proatoms(hp_part *p, const double *coefs_new, const double *coefs_old, double *min_proatom,
                            double *msd) {
  if (!p || !coefs_new || !coefs_old) return fail(HP_E_ARG, "NULL argument");
  if (p->scheme != HP_SCHEME_LISA) return fail(HP_E_UNSUPP, "the global terms need a linear-coefficient basis (LISA)");
  HP_HIP(hipSetDevice(p->device));
  if (int rc = glisa_setup(p)) return rc;
  hipStream_t s = p->stream;
  const int M = p->gl_M;
  HP_HIP(hipMemcpyAsync(p->gl_coef, coefs_new, sizeof(double) * M, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(p->gl_coef + M, coefs_old, sizeof(double) * M, hipMemcpyHostToDevice, s));
  DevFn f = p->fn();
  f.alpha = p->fn_alpha;
  const dim3 g((unsigned)p->gl_nchunk, (unsigned)p->natom), b(256);
  double *part = p->gl_pro, *out = p->gl_pro + 2 * (size_t)p->natom * p->gl_nchunk;
#define HP_GL_PRO(FK)                                                                                             \
  hipLaunchKernelGGL((k_glisa_proatoms<FK>), g, b, 0, s, p->mol(), f, p->fn_norm, p->gl_coef, p->gl_coef + M, part)
  if (p->fnkind == HP_FN_SLATER)
    HP_GL_PRO(HP_FN_SLATER);
  else if (p->fnkind == HP_FN_GAUSS)
    HP_GL_PRO(HP_FN_GAUSS);
  else
    HP_GL_PRO(HP_FN_GENERAL);
#undef HP_GL_PRO
  HP_LAUNCHED();
  hipLaunchKernelGGL(k_glisa_proatoms_reduce, dim3((p->natom + 63) / 64), dim3(64), 0, s, p->natom, p->gl_nchunk, part,
                     out);
  HP_LAUNCHED();
  std::vector<double> h(2 * (size_t)p->natom);
  HP_HIP(hipMemcpyAsync(h.data(), out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  HP_HIP(hipStreamSynchronize(s));
  for (int a = 0; a < p->natom; ++a) {
    if (min_proatom) min_proatom[a] = h[2 * a];
    if (msd) msd[a] = h[2 * a + 1];
  }
  return 0;
}

int hp_part_populations(hp_part *p, double *pop) {
  if (!p || !pop) return fail(HP_E_ARG, "NULL argument");
  if (p->grid_type != 2 && !p->tabulated()) return fail(HP_E_UNSUPP, "molecular-grid populations need grid_type 2 or 3");
  if (p->niter == 0 && !p->weights_ready) return fail(HP_E_ARG, "run a weights pass first");
  HP_HIP(hipSetDevice(p->device));
  if (p->tabulated() && !p->only_molgrid) {
    // atom-grid populations of the last tabulated weights pass, summed per
    // atom over its shells in order (grid_type 2 keeps the weights on the
    // molecular grid but integrates on the atom grids, core/base.py:293-299)
    std::vector<double> part(p->nshell);
    HP_HIP(hipMemcpyAsync(part.data(), p->pop_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, p->stream));
    HP_HIP(hipStreamSynchronize(p->stream));
    for (int a = 0; a < p->natom; ++a) {
      double v = 0.0;
      for (int si = p->h_atom_shell_off[a]; si < p->h_atom_shell_off[a + 1]; ++si) v += part[si];
      pop[a] = v;
    }
    return 0;
  }
  double *part = nullptr, *out = nullptr;
  const size_t npart = moments_partials(p, true);
  int rc = 0;
  if ((rc = dalloc(&part, npart * HP_NMOM)) || (rc = dalloc(&out, (size_t)p->natom * HP_NMOM))) {
    if (part) (void)hipFree(part);
    return rc;
  }
  rc = launch_moments(p, part, out, true);
  std::vector<double> h((size_t)p->natom * HP_NMOM);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(h.data(), out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = fail(HP_E_HIP, std::string("populations: ") + hipGetErrorString(e));
  }
  (void)hipFree(part);
  (void)hipFree(out);
  if (rc) return rc;
  for (int a = 0; a < p->natom; ++a) pop[a] = h[(size_t)a * HP_NMOM];
  return 0;
}

int hp_part_kernel_time(hp_part *p, int kernel, double *total_ms, int64_t *launches) {
  if (!p || kernel < 0 || kernel > HP_KERNEL_GLISA_HESS) return fail(HP_E_ARG, "bad argument");
  if (total_ms) *total_ms = p->k_ms[kernel];
  if (launches) *launches = p->k_n[kernel];
  return 0;
}

int hp_part_get_work(hp_part *p, hp_part_work *out) {
  if (!p || !out) return fail(HP_E_ARG, "NULL argument");
  HP_HIP(hipSetDevice(p->device));
  out->points = p->npoint;
  out->evals_per_pass = p->evals_per_pass;
  // tabulated weights pass: x, y, z, rho and three weights per point (and
  // the promolecule on the molecular grid), three sums per shell
  out->bytes_per_pass = p->tabulated() ? (56.0 + (p->promol ? 8.0 : 0.0)) * (double)p->npoint + 24.0 * (double)p->nshell
                                       : 48.0 * (double)p->npoint + 16.0 * (double)p->nshell;
  out->hessian_passes = 0;
  out->hessian_flops = 0.0;
  if (int rc = hp::read_eval_counters(p->evals_done, &out->evals_done_fp32, &out->evals_done_fp64)) return rc;
  if (p->gt2_nhess) {
    std::vector<int> nh(p->natom);
    HP_HIP(hipMemcpy(nh.data(), p->gt2_nhess, sizeof(int) * p->natom, hipMemcpyDeviceToHost));
    std::vector<int> off(p->natom + 1);
    HP_HIP(hipMemcpy(off.data(), p->fn_off, sizeof(int) * (p->natom + 1), hipMemcpyDeviceToHost));
    for (int a = 0; a < p->natom; ++a) {
      const double K = off[a + 1] - off[a];
      out->hessian_passes += nh[a];
      out->hessian_flops += nh[a] * 2.0 * K * K * (double)p->npoint;
    }
  }
  return 0;
}

}  // extern "C"

// ---- grid utilities ----------------------------------------------------------

namespace {
// Device buffers and a stream of the calling host thread for the grid
// utilities: they grow to the largest molecule seen and are kept until the
// process ends, so that building a request's inputs while a batch serves
// allocates and frees nothing (hipFree waits for the whole device) and waits
// only for its own stream.
struct GridScratch {
  int device = -1;
  hipStream_t stream = nullptr;
  double *buf[8] = {};
  size_t cap[8] = {};
};

thread_local GridScratch t_grid;

int grid_scratch(int device, GridScratch **out) {
  GridScratch &g = t_grid;
  if (g.device != device) {
    // a thread that changes devices leaves its buffers of the previous one
    g = GridScratch();
    HP_HIP(hipSetDevice(device));
    HP_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    g.device = device;
  } else {
    HP_HIP(hipSetDevice(device));
  }
  *out = &g;
  return 0;
}

// Buffer k of at least n bytes.
int grid_buf(GridScratch *g, int k, size_t n, void **out) {
  if (g->cap[k] < n) {
    const size_t want = std::max(n, g->cap[k] + g->cap[k] / 2);
    if (g->buf[k]) HP_HIP(hipFree(g->buf[k]));
    g->buf[k] = nullptr;
    g->cap[k] = 0;
    HP_HIP(hipMalloc((void **)&g->buf[k], want));
    g->cap[k] = want;
  }
  *out = g->buf[k];
  return 0;
}

// Atom centres and points as SoA device columns on the thread's stream
// (buffers 0-1: centres, points).
int upload_points(GridScratch *g, int natom, const double *coords, int64_t npoint, const double *points, double **c,
                  double **p) {
  std::vector<double> a3(3 * (size_t)natom);
  for (int a = 0; a < natom; ++a)
    for (int d = 0; d < 3; ++d) a3[(size_t)d * natom + a] = coords[3 * a + d];
  std::vector<double> p3(3 * (size_t)npoint);
  for (int64_t i = 0; i < npoint; ++i)
    for (int d = 0; d < 3; ++d) p3[(size_t)d * npoint + i] = points[3 * i + d];
  void *dc = nullptr, *dp = nullptr;
  int rc = 0;
  if ((rc = grid_buf(g, 0, a3.size() * sizeof(double), &dc)) || (rc = grid_buf(g, 1, p3.size() * sizeof(double), &dp)))
    return rc;
  HP_HIP(hipMemcpyAsync(dc, a3.data(), a3.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HP_HIP(hipMemcpyAsync(dp, p3.data(), p3.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  // the host staging vectors go out of scope: wait for the copies
  HP_HIP(hipStreamSynchronize(g->stream));
  *c = (double *)dc;
  *p = (double *)dp;
  return 0;
}
}  // namespace

extern "C" {

int hp_becke_weights(int device, int natom, const double *coordinates, const double *radii, int64_t npoint,
                     const double *points, const int64_t *atom_offsets, double *out) {
  g_err.clear();
  if (natom <= 0 || npoint <= 0 || !coordinates || !radii || !points || !atom_offsets || !out)
    return fail(HP_E_ARG, "bad argument");
  if (atom_offsets[0] != 0 || atom_offsets[natom] != npoint) return fail(HP_E_ARG, "atom_offsets must span the points");
  for (int a = 0; a < natom; ++a)
    if (atom_offsets[a + 1] < atom_offsets[a]) return fail(HP_E_ARG, "atom_offsets must be non-decreasing");
  // size adjustment a_ij from the radius ratio chi = R_i / R_j, |a_ij| <= 1/2
  std::vector<double> tab(2 * (size_t)natom * natom, 0.0);
  double *adj = tab.data(), *invr = tab.data() + (size_t)natom * natom;
  for (int i = 0; i < natom; ++i)
    for (int j = 0; j < natom; ++j) {
      if (i == j) continue;
      const double chi = radii[i] / radii[j];
      const double u = (chi - 1.0) / (chi + 1.0);
      const double den = u * u - 1.0;
      double aij = den != 0.0 ? u / den : 0.0;
      aij = std::min(0.5, std::max(-0.5, aij));
      adj[(size_t)i * natom + j] = aij;
      const double dx = coordinates[3 * i] - coordinates[3 * j], dy = coordinates[3 * i + 1] - coordinates[3 * j + 1],
                   dz = coordinates[3 * i + 2] - coordinates[3 * j + 2];
      const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (d == 0.0) return fail(HP_E_ARG, "two atoms at the same position");
      invr[(size_t)i * natom + j] = 1.0 / d;
    }
  // per atom o: every atom by distance from o (o first), ties by index
  std::vector<int> ord((size_t)natom * natom);
  {
    std::vector<double> d2(natom);
    for (int o = 0; o < natom; ++o) {
      for (int j = 0; j < natom; ++j) {
        const double dx = coordinates[3 * o] - coordinates[3 * j], dy = coordinates[3 * o + 1] - coordinates[3 * j + 1],
                     dz = coordinates[3 * o + 2] - coordinates[3 * j + 2];
        d2[j] = dx * dx + dy * dy + dz * dz;
      }
      int *row = ord.data() + (size_t)o * natom;
      for (int j = 0; j < natom; ++j) row[j] = j;
      std::sort(row, row + natom, [&](int u, int v) { return d2[u] < d2[v] || (d2[u] == d2[v] && u < v); });
      if (row[0] != o) return fail(HP_E_ARG, "two atoms at the same position");
    }
  }
  GridScratch *g = nullptr;
  double *c = nullptr, *p = nullptr;
  void *dtab = nullptr, *doff = nullptr, *dout = nullptr, *dord = nullptr;
  int rc = grid_scratch(device, &g);
  if (!rc) rc = upload_points(g, natom, coordinates, npoint, points, &c, &p);
  if (!rc) rc = grid_buf(g, 2, tab.size() * sizeof(double), &dtab);
  if (!rc) rc = grid_buf(g, 3, (natom + 1) * sizeof(int64_t), &doff);
  if (!rc) rc = grid_buf(g, 4, npoint * sizeof(double), &dout);
  if (!rc) rc = grid_buf(g, 7, ord.size() * sizeof(int), &dord);
  if (rc) return rc;
  HP_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HP_HIP(hipMemcpyAsync(doff, atom_offsets, (natom + 1) * sizeof(int64_t), hipMemcpyHostToDevice, g->stream));
  HP_HIP(hipMemcpyAsync(dord, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice, g->stream));
  const double *dadj = (const double *)dtab, *dinvr = (const double *)dtab + (size_t)natom * natom;
  hipLaunchKernelGGL(k_becke, dim3((unsigned)((npoint + 255) / 256)), dim3(256), 0, g->stream, natom, c, c + natom,
                     c + 2 * natom, dadj, dinvr, (const int *)dord, npoint, p, p + npoint, p + 2 * npoint,
                     (const int64_t *)doff, 3, (double *)dout);
  HP_LAUNCHED();
  HP_HIP(hipMemcpyAsync(out, dout, sizeof(double) * npoint, hipMemcpyDeviceToHost, g->stream));
  HP_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int hp_slater_density(int device, int natom, const double *coordinates, const int32_t *fn_offsets,
                      const double *coef, const double *alpha, int64_t npoint, const double *points, double *out) {
  g_err.clear();
  if (natom <= 0 || npoint <= 0 || !coordinates || !fn_offsets || !coef || !alpha || !points || !out)
    return fail(HP_E_ARG, "bad argument");
  const int nfn = fn_offsets[natom];
  if (fn_offsets[0] != 0 || nfn <= 0) return fail(HP_E_ARG, "bad fn_offsets");
  for (int a = 0; a < natom; ++a)
    if (fn_offsets[a + 1] < fn_offsets[a]) return fail(HP_E_ARG, "fn_offsets must be non-decreasing");
  std::vector<double> tab(2 * (size_t)nfn);
  std::copy(coef, coef + nfn, tab.begin());
  std::copy(alpha, alpha + nfn, tab.begin() + nfn);
  // the workgroups' bounds need every term to decrease with distance
  bool cull = true;
  for (int f = 0; f < nfn; ++f) cull = cull && coef[f] >= 0.0 && alpha[f] >= 0.0;
  GridScratch *g = nullptr;
  double *c = nullptr, *p = nullptr;
  void *dtab = nullptr, *doff = nullptr, *dout = nullptr;
  int rc = grid_scratch(device, &g);
  if (!rc) rc = upload_points(g, natom, coordinates, npoint, points, &c, &p);
  if (!rc) rc = grid_buf(g, 5, tab.size() * sizeof(double), &dtab);
  if (!rc) rc = grid_buf(g, 6, (natom + 1) * sizeof(int32_t), &doff);
  if (!rc) rc = grid_buf(g, 4, npoint * sizeof(double), &dout);
  if (rc) return rc;
  HP_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HP_HIP(hipMemcpyAsync(doff, fn_offsets, (natom + 1) * sizeof(int32_t), hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(k_slater_sum, dim3((unsigned)((npoint + 255) / 256)), dim3(256), 0, g->stream, natom, c, c + natom,
                     c + 2 * natom, (const int *)doff, (const double *)dtab, (const double *)dtab + nfn, npoint, p,
                     p + npoint, p + 2 * npoint, cull ? 1 : 0, (double *)dout);
  HP_LAUNCHED();
  HP_HIP(hipMemcpyAsync(out, dout, sizeof(double) * npoint, hipMemcpyDeviceToHost, g->stream));
  HP_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int hp_stream_copy(int device, int64_t nbytes, int reps, double *gbps, double *ms_per_copy) {
  g_err.clear();
  if (nbytes < 16 || reps < 1 || !gbps || !ms_per_copy) return fail(HP_E_ARG, "bad argument");
  HP_HIP(hipSetDevice(device));
  const int64_t n2 = nbytes / 16;
  double2 *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = dalloc(&src, n2);
  if (!rc) rc = dalloc(&dst, n2);
  hipError_t e = hipSuccess;
  if (!rc) {
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    const unsigned grid = (unsigned)std::min<int64_t>((n2 + 255) / 256, (int64_t)ncu * 16);
    e = hipMemset(src, 0, n2 * sizeof(double2));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // one untimed copy, then reps timed ones
    for (int r = 0; e == hipSuccess && r <= reps; ++r) {
      if (r == 1) e = hipEventRecord(e0, nullptr);
      hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, nullptr, src, dst, n2);
      if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) {
      rc = fail(HP_E_HIP, std::string("copy: ") + hipGetErrorString(e));
    } else {
      *ms_per_copy = ms / reps;
      *gbps = 2.0 * (double)(n2 * 16) / (*ms_per_copy * 1e-3) / 1e9;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

int hp_exp_rate(int device, int precision, double *exps_per_s) {
  g_err.clear();
  if (!exps_per_s || (precision != HP_PREC_FP64 && precision != HP_PREC_MIXED)) return fail(HP_E_ARG, "bad argument");
  HP_HIP(hipSetDevice(device));
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  const int block = 256, grid = ncu * 16, iters = 256;
  double *out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = dalloc(&out, (size_t)grid * block);
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int r = 0; e == hipSuccess && r < 2; ++r) {  // the first launch is untimed
      if (r == 1) e = hipEventRecord(e0, nullptr);
      if (precision == HP_PREC_FP64)
        hipLaunchKernelGGL(k_exp_rate<double>, dim3(grid), dim3(block), 0, nullptr, iters, out);
      else
        hipLaunchKernelGGL(k_exp_rate<float>, dim3(grid), dim3(block), 0, nullptr, iters, out);
      if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
      rc = fail(HP_E_HIP, std::string("exp rate: ") + hipGetErrorString(e));
    else
      *exps_per_s = (double)grid * block * iters * HP_EXP_RATE_CHAINS / (ms * 1e-3);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  return rc;
}

}  // extern "C"
