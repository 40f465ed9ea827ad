// Continuous batching of stockholder partitionings on one GPU.
//
// Many molecules are resident at once in one device arena: their points,
// radial shells, atoms and pro-atom functions occupy ranges of shared
// columns, so the kernels of the single-molecule path run over all of them
// with one launch per kernel variant.  A launch's blocks (or waves) take
// their (shell or atom, slot) pair from a work list; a molecule's slot record
// (HpSlot) holds its outer-loop state, advanced on the device by
// k_batch_finalize with the same stop rules as hp_part_run.
//
// Host side of a step:
//   1. wait for the previous step's ticks;
//   2. retire molecules that stopped (results copied out, ranges freed) and
//      move molecules whose FP32 stage ended to the FP64 lists;
//   3. let molecules whose upload has finished join;
//   4. rebuild the work lists if the set of molecules changed;
//   5. enqueue ticks_per_step ticks (or replay them from a captured graph);
//   6. admit queued requests into free ranges: their arrays are copied on a
//      second stream while the GPU runs the ticks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "hortonpart.h"
#include "hp_device.h"
#include "hp_host.h"

namespace {

using hp::dalloc;
using hp::fail;

// First-fit allocator of index ranges in [0, capacity).
class RangeAlloc {
 public:
  void init(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  // offset of a free range of n entries, or -1
  int64_t alloc(int64_t n) {
    if (n <= 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int64_t, int64_t> free_;
};

enum SlotState { SLOT_FREE = 0, SLOT_UPLOADING, SLOT_ACTIVE };

constexpr int kMaxGroups = 4;  // groups of molecules whose ticks interleave (one stream each)

// Default group count.  Each group's stream needs a hardware queue of its
// own for the groups to overlap; streams beyond the process's queues share
// one and run in turn.  HIP gives a process GPU_MAX_HW_QUEUES queues (4 when
// unset), and the engine also uses an upload stream, so three groups need
// at least 8 queues; two groups otherwise.
int default_groups() {
  const char *e = std::getenv("GPU_MAX_HW_QUEUES");
  const int q = e ? std::atoi(e) : 4;
  return q >= 8 ? 3 : 2;
}

struct Request {
  int64_t ticket = 0;
  hp_molecule_desc md{};
  hp_scheme_desc sd{};
  const double *propars0 = nullptr;
  double *propars_out = nullptr, *charges_out = nullptr;
  int precision = HP_PREC_FP64, n_polish = 1;
  // sizes
  int A = 0, S = 0, F = 0, npar = 0, maxL = 0, maxR = 0, maxK = 0;
  int64_t N = 0, Q = 0;
  int scheme01 = 0;  // 0 MBIS, 1 GISA
  // placement in the arena
  int slot = -1, a0 = 0, s0 = 0, f0 = 0;
  int64_t n0 = 0, q0 = 0;
  int stage = HP_RUN_FP64;  // run value the host believes the slot has
  int corr_period = 0;      // FP32 stage: FP64 refresh period of the correction
  // weights pass: workgroup shape, and the first radial shell (local) of each
  // workgroup; both depend on the molecule alone, so that its sums are the
  // same in any mix
  hp::WeightsShape wshape{1, 64, 1};
  std::vector<int> grp_first;
};

// One kernel launch of a tick.
struct Launch {
  int kind;  // 0 weights, 1 MBIS inner, 2 GISA QP
  int64_t off, len;
  // weights
  int fp64, fk, bs, pt, tile;
  size_t lds;
  // per-atom kernels
  int km, block, ppl;
};

}  // namespace

struct hp_batch {
  int device = 0;
  hp_batch_config cfg{};
  hipStream_t cs = nullptr, us = nullptr;  // compute, upload
  // points
  double *px = nullptr, *py = nullptr, *pz = nullptr, *rho = nullptr, *wmol = nullptr, *wsph = nullptr;
  // shells
  int *shell_atom = nullptr;
  int64_t *shell_off = nullptr;
  double *shell_r = nullptr, *shell_w = nullptr, *rhobar = nullptr, *ent_part = nullptr, *corr = nullptr;
  // atoms
  double *cx = nullptr, *cy = nullptr, *cz = nullptr, *pseudo = nullptr, *charges = nullptr, *change_part = nullptr;
  int *atom_shell_off = nullptr, *fn_off = nullptr, *inner_niter = nullptr, *status = nullptr, *status_all = nullptr;
  int64_t *qp_off = nullptr;
  // functions
  double *coef = nullptr, *alpha = nullptr, *fn_norm = nullptr, *fn_alpha = nullptr, *fn_order = nullptr;
  double *par_l = nullptr, *old_l = nullptr, *par_m = nullptr, *old_m = nullptr;
  double *qpP = nullptr;
  unsigned long long *evals_done = nullptr;  // [FP32, FP64] evaluations performed
  // MBIS inner solves spread over launches: per atom the iterations done (-1:
  // none in progress), per radial shell the previous pro-atom value
  int *ist_it = nullptr;
  double *ist_oldpro = nullptr;
  // slots
  HpSlot *slots = nullptr, *h_slots = nullptr, *h_slot_init = nullptr;
  std::vector<int> slot_state;
  std::vector<Request *> slot_req;
  std::vector<hipEvent_t> up_ev;
  std::vector<int> free_slots;
  // staging of one admission (upload stream)
  double *st_points = nullptr, *st_coords = nullptr;
  int *st_shell_atom = nullptr, *st_fn_off = nullptr;
  int64_t *st_shell_off = nullptr;
  // work lists
  int2 *d_work = nullptr, *h_work = nullptr;
  int *d_slot_list = nullptr, *h_slot_list = nullptr;
  int64_t work_cap = 0;
  int nslot_list = 0;
  // groups of molecules, group g launched on gs[g] (gs[0] is cs)
  std::vector<Launch> plan[kMaxGroups];
  int nslot[kMaxGroups] = {}, slot_off[kMaxGroups] = {};
  hipStream_t gs[kMaxGroups] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxGroups] = {}, ev_step[2] = {nullptr, nullptr};
  std::vector<int> pending_fp64;  // slots to flip to FP64 before the next ticks
  // graph of one step's ticks
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // allocators and queues
  RangeAlloc pts, atoms, shells, fns, qps;
  std::deque<Request *> queue;
  std::deque<hp_batch_result> done;
  int64_t next_ticket = 1;
  bool ticks_in_flight = false;
  // per tick of a step and group: events at its start, after k_batch_begin,
  // after the weights launches, after the per-atom launches and at its end
  std::vector<hipEvent_t> tev;
  std::vector<int> last_niter;  // per slot: iterations already counted
  hp_batch_stats stats{};

  DevMol mol(int64_t work_off, int natom_list) const {
    DevMol m;
    m.natom = natom_list;
    m.nshell = 0;
    m.npoint = cfg.point_capacity;
    m.cx = cx;
    m.cy = cy;
    m.cz = cz;
    m.px = px;
    m.py = py;
    m.pz = pz;
    m.rho = rho;
    m.wmol = wmol;
    m.wat = nullptr;
    m.wsph = wsph;
    m.shell_atom = shell_atom;
    m.shell_off = shell_off;
    m.shell_r = shell_r;
    m.shell_w = shell_w;
    m.atom_shell_off = atom_shell_off;
    m.pseudo = pseudo;
    m.work = d_work + work_off;
    m.slots = slots;
    m.grp_first = nullptr;  // the work list names each group's first shell
    m.wgroup = 1;           // per molecule: HpSlot::wgroup
    return m;
  }
  DevFn fn() const {
    DevFn f;
    f.fn_off = fn_off;
    f.coef = coef;
    f.alpha = alpha;
    f.order = fn_order;
    f.nb_list = nullptr;
    f.shell_nb = nullptr;
    f.nb_off = nullptr;
    f.rc = 0.0;
    f.tail_cull = 1;  // MBIS and GISA only; per request through HpSlot::full_sum
    f.evals_done = evals_done;
    return f;
  }
};

namespace {

void free_batch(hp_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void *bufs[] = {b->px,         b->py,          b->pz,          b->rho,         b->wmol,      b->wsph,
                  b->shell_atom, b->shell_off,   b->shell_r,     b->shell_w,     b->rhobar,    b->ent_part,
                  b->corr,
                  b->cx,         b->cy,          b->cz,          b->pseudo,      b->charges,   b->change_part,
                  b->atom_shell_off, b->fn_off,  b->inner_niter, b->status,      b->status_all, b->qp_off,
                  b->coef,       b->alpha,       b->fn_norm,     b->fn_alpha,    b->fn_order,  b->par_l,
                  b->old_l,      b->par_m,       b->old_m,       b->qpP,         b->slots,     b->st_points,
                  b->st_coords,  b->st_shell_atom, b->st_fn_off, b->st_shell_off, b->d_work,  b->d_slot_list,
                  b->evals_done,  b->ist_it,     b->ist_oldpro};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (b->h_slots) (void)hipHostFree(b->h_slots);
  if (b->h_slot_init) (void)hipHostFree(b->h_slot_init);
  if (b->h_work) (void)hipHostFree(b->h_work);
  if (b->h_slot_list) (void)hipHostFree(b->h_slot_list);
  for (auto e : b->up_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto e : b->tev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {b->ev_fork, b->ev_step[0], b->ev_step[1]})
    if (e) (void)hipEventDestroy(e);
  for (int g = 1; g < kMaxGroups; ++g) {
    if (b->ev_join[g]) (void)hipEventDestroy(b->ev_join[g]);
    if (b->gs[g]) (void)hipStreamDestroy(b->gs[g]);
  }
  for (auto *r : b->slot_req) delete r;
  for (auto *r : b->queue) delete r;
  if (b->cs) (void)hipStreamDestroy(b->cs);
  if (b->us) (void)hipStreamDestroy(b->us);
  delete b;
}

// ---- validation of one request ----------------------------------------------

int inspect(const hp_batch *b, Request *r) {
  const hp_molecule_desc *md = &r->md;
  const hp_scheme_desc *sd = &r->sd;
  const int A = md->natom, S = md->nshell;
  const int64_t N = md->npoint;
  if (A <= 0 || S <= 0 || N <= 0) return fail(HP_E_ARG, "natom, nshell and npoint must be positive");
  if (!md->coordinates || !md->pseudo_numbers || !md->points || !md->density || !md->mol_weights ||
      !md->sph_weights || !md->shell_atom || !md->shell_offsets || !md->shell_radius || !md->shell_rweight)
    return fail(HP_E_ARG, "molecule description has a NULL array");
  if (sd->grid_type != 1) return fail(HP_E_UNSUPP, "a batch runs grid_type 1 only");
  if (!(sd->scheme == HP_SCHEME_MBIS && sd->solver == HP_SOLVER_MBIS) &&
      !(sd->scheme == HP_SCHEME_GISA && sd->solver == HP_SOLVER_QP))
    return fail(HP_E_UNSUPP, "a batch runs MBIS and GISA (QP) requests");
  if (sd->cutoff_radius > 0.0) return fail(HP_E_UNSUPP, "a batch has no cut-off radius");
  // the settings the kernels take per launch must be the batch's
  const bool same_mbis = sd->scheme != HP_SCHEME_MBIS ||
                         (sd->inner_threshold == b->cfg.inner_threshold && sd->inner_maxiter == b->cfg.inner_maxiter);
  const bool same_qp = sd->scheme != HP_SCHEME_GISA || (sd->negative_cutoff == b->cfg.negative_cutoff &&
                                                        sd->population_cutoff == b->cfg.population_cutoff);
  if (!same_mbis || !same_qp || sd->density_cutoff != b->cfg.density_cutoff)
    return fail(HP_E_ARG, "inner-solver settings differ from the batch's");
  if (sd->maxiter < 1) return fail(HP_E_ARG, "maxiter must be positive");
  if (r->precision != HP_PREC_FP64 && r->precision != HP_PREC_MIXED) return fail(HP_E_ARG, "unknown precision");
  if (r->n_polish < 0) return fail(HP_E_ARG, "n_polish must be >= 0");
  if (!sd->fn_offsets || !r->propars0 || !r->propars_out || !r->charges_out) return fail(HP_E_ARG, "NULL array");
  const bool mbis = sd->scheme == HP_SCHEME_MBIS;
  if (!mbis && (!sd->fn_alpha || !sd->fn_order || !sd->fn_norm || !sd->qp_matrix))
    return fail(HP_E_ARG, "basis arrays are NULL");
  // grid structure: shells sorted by atom, every atom at least one shell
  std::vector<int> nsh(A, 0);
  int prev = -1;
  for (int s = 0; s < S; ++s) {
    const int a = md->shell_atom[s];
    if (a < 0 || a >= A || a < prev) return fail(HP_E_ARG, "shell_atom must be sorted atom indices");
    nsh[a] += 1;
    prev = a;
  }
  int maxR = 0;
  for (int a = 0; a < A; ++a) {
    if (nsh[a] == 0) return fail(HP_E_ARG, "every atom needs at least one radial shell");
    maxR = std::max(maxR, nsh[a]);
  }
  if (md->shell_offsets[0] != 0 || md->shell_offsets[S] != N) return fail(HP_E_ARG, "shell_offsets must span the grid");
  int maxL = 0;
  for (int s = 0; s < S; ++s) {
    const int64_t L = md->shell_offsets[s + 1] - md->shell_offsets[s];
    if (L <= 0) return fail(HP_E_ARG, "empty radial shell");
    maxL = (int)std::max<int64_t>(maxL, L);
  }
  if (maxL > 1024) return fail(HP_E_UNSUPP, "more than 1024 points in one radial shell");
  if (maxR > 512) return fail(HP_E_UNSUPP, "more than 512 radial points per atom");
  if (sd->fn_offsets[0] != 0) return fail(HP_E_ARG, "fn_offsets[0] must be 0");
  int maxK = 0;
  int64_t Q = 0;
  for (int a = 0; a < A; ++a) {
    const int K = sd->fn_offsets[a + 1] - sd->fn_offsets[a];
    if (K <= 0) return fail(HP_E_ARG, "every atom needs at least one function");
    maxK = std::max(maxK, K);
    Q += (int64_t)K * K;
  }
  if (mbis && maxK > 7) return fail(HP_E_UNSUPP, "MBIS supports at most 7 shells per atom");
  if (maxK > HP_KMAX) return fail(HP_E_UNSUPP, "too many functions per atom");
  const int F = sd->fn_offsets[A];
  if (!mbis) {
    for (int f = 0; f < F; ++f)
      if (sd->fn_order[f] != 2.0 || !(sd->fn_alpha[f] >= 0.0))
        return fail(HP_E_UNSUPP, "GISA in a batch takes Gaussian functions (order 2) with exponents >= 0");
  }
  if (N > b->cfg.max_molecule_points) return fail(HP_E_ARG, "molecule has more points than max_molecule_points");
  if (N > b->cfg.point_capacity || A + 1 > b->cfg.atom_capacity || S + 1 > b->cfg.shell_capacity ||
      F > b->cfg.function_capacity || (!mbis && Q > b->cfg.qp_capacity))
    return fail(HP_E_ARG, "molecule is larger than the batch arena");
  r->A = A;
  r->S = S;
  r->F = F;
  r->N = N;
  r->Q = mbis ? 0 : Q;
  r->npar = mbis ? 2 * F : F;
  r->maxL = maxL;
  r->maxR = maxR;
  r->maxK = maxK;
  r->scheme01 = mbis ? 0 : 1;
  r->wshape = hp::weights_shape(maxL);
  std::vector<int> aso(A + 1, 0);
  for (int a = 0; a < A; ++a) aso[a + 1] = aso[a] + nsh[a];
  r->grp_first = hp::weights_groups(aso, r->wshape.g);
  return 0;
}

// ---- admission ----------------------------------------------------------------

// Reserve arena ranges and a slot; false if the request does not fit now.
bool place(hp_batch *b, Request *r) {
  if (b->free_slots.empty()) return false;
  const int64_t n0 = b->pts.alloc(r->N);
  if (n0 < 0) return false;
  const int64_t a0 = b->atoms.alloc(r->A + 1);
  if (a0 < 0) {
    b->pts.release(n0, r->N);
    return false;
  }
  const int64_t s0 = b->shells.alloc(r->S + 1);
  if (s0 < 0) {
    b->pts.release(n0, r->N);
    b->atoms.release(a0, r->A + 1);
    return false;
  }
  const int64_t f0 = b->fns.alloc(r->F);
  if (f0 < 0) {
    b->pts.release(n0, r->N);
    b->atoms.release(a0, r->A + 1);
    b->shells.release(s0, r->S + 1);
    return false;
  }
  const int64_t q0 = b->qps.alloc(r->Q);
  if (q0 < 0) {
    b->pts.release(n0, r->N);
    b->atoms.release(a0, r->A + 1);
    b->shells.release(s0, r->S + 1);
    b->fns.release(f0, r->F);
    return false;
  }
  r->n0 = n0;
  r->a0 = (int)a0;
  r->s0 = (int)s0;
  r->f0 = (int)f0;
  r->q0 = q0;
  r->slot = b->free_slots.back();
  b->free_slots.pop_back();
  return true;
}

void unplace(hp_batch *b, Request *r) {
  b->pts.release(r->n0, r->N);
  b->atoms.release(r->a0, r->A + 1);
  b->shells.release(r->s0, r->S + 1);
  b->fns.release(r->f0, r->F);
  b->qps.release(r->q0, r->Q);
  b->free_slots.push_back(r->slot);
  r->slot = -1;
}

// Copy a placed request's arrays into its ranges (upload stream).
int upload(hp_batch *b, Request *r) {
  hipStream_t s = b->us;
  const hp_molecule_desc *md = &r->md;
  const hp_scheme_desc *sd = &r->sd;
  const size_t d8 = sizeof(double);
  const int64_t N = r->N;
  HP_HIP(hipMemcpyAsync(b->st_points, md->points, 3 * N * d8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_batch_points, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, b->st_points, N, b->px + r->n0,
                     b->py + r->n0, b->pz + r->n0);
  HP_LAUNCHED();
  HP_HIP(hipMemcpyAsync(b->rho + r->n0, md->density, N * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->wmol + r->n0, md->mol_weights, N * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->wsph + r->n0, md->sph_weights, N * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->shell_r + r->s0, md->shell_radius, r->S * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->shell_w + r->s0, md->shell_rweight, r->S * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->pseudo + r->a0, md->pseudo_numbers, r->A * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->st_coords, md->coordinates, 3 * r->A * d8, hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->st_shell_atom, md->shell_atom, r->S * sizeof(int), hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->st_shell_off, md->shell_offsets, (r->S + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  HP_HIP(hipMemcpyAsync(b->st_fn_off, sd->fn_offsets, (r->A + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  if (r->scheme01 == 0) {
    HP_HIP(hipMemcpyAsync(b->par_m + 2 * (int64_t)r->f0, r->propars0, r->npar * d8, hipMemcpyHostToDevice, s));
  } else {
    HP_HIP(hipMemcpyAsync(b->par_l + r->f0, r->propars0, r->npar * d8, hipMemcpyHostToDevice, s));
    HP_HIP(hipMemcpyAsync(b->fn_norm + r->f0, sd->fn_norm, r->F * d8, hipMemcpyHostToDevice, s));
    HP_HIP(hipMemcpyAsync(b->fn_alpha + r->f0, sd->fn_alpha, r->F * d8, hipMemcpyHostToDevice, s));
    HP_HIP(hipMemcpyAsync(b->fn_order + r->f0, sd->fn_order, r->F * d8, hipMemcpyHostToDevice, s));
    HP_HIP(hipMemcpyAsync(b->qpP + r->q0, sd->qp_matrix, r->Q * d8, hipMemcpyHostToDevice, s));
  }
  HpAdmit ad;
  ad.a0 = r->a0;
  ad.na = r->A;
  ad.s0 = r->s0;
  ad.ns = r->S;
  ad.f0 = r->f0;
  ad.n0 = r->n0;
  ad.q0 = r->q0;
  ad.has_qp = r->scheme01 == 1;
  hipLaunchKernelGGL(k_batch_admit, dim3(1), dim3(256), 0, s, ad, b->st_coords, b->st_shell_atom, b->st_shell_off,
                     b->st_fn_off, b->cx, b->cy, b->cz, b->shell_atom, b->shell_off, b->atom_shell_off, b->fn_off,
                     b->qp_off, b->inner_niter, b->status, b->status_all, b->charges, b->change_part, b->ist_it);
  HP_LAUNCHED();
  // the slot record: idle until the host lets the molecule join the lists
  HpSlot &h = b->h_slot_init[r->slot];
  std::memset(&h, 0, sizeof(h));
  h.run = HP_RUN_IDLE;
  h.event = HP_EV_RUNNING;
  h.scheme = r->scheme01;
  h.a0 = r->a0;
  h.na = r->A;
  h.s0 = r->s0;
  h.ns = r->S;
  h.f0 = r->f0;
  h.nf = r->F;
  h.niter = 0;
  h.min_stop = 1;
  h.max_iter = r->sd.maxiter;
  h.n_polish = r->n_polish;
  h.maxiter = r->sd.maxiter;
  h.threshold = r->sd.threshold;
  const double sw = r->sd.mixed_switch > 0.0 ? r->sd.mixed_switch : 0.0;
  h.thr = r->precision == HP_PREC_MIXED ? std::max(r->sd.threshold * HP_MIXED_MARGIN, sw) : r->sd.threshold;
  h.corr_period = r->precision == HP_PREC_MIXED && r->sd.mixed_refresh > 0 ? r->sd.mixed_refresh : 0;
  h.refresh = 0;
  // tail skipping needs non-negative coefficients: MBIS and GISA updates keep
  // them so, from non-negative initial parameters
  bool nonneg = true;
  for (int64_t k = 0; k < r->npar; ++k) nonneg = nonneg && r->propars0[k] >= 0.0;
  h.full_sum = (r->sd.full_sum || !nonneg) ? 1 : 0;
  h.wgroup = r->wshape.g;
  r->stage = r->precision == HP_PREC_MIXED ? HP_RUN_FP32 : HP_RUN_FP64;
  r->corr_period = h.corr_period;
  HP_HIP(hipMemcpyAsync(b->slots + r->slot, &h, sizeof(HpSlot), hipMemcpyHostToDevice, s));
  HP_HIP(hipEventRecord(b->up_ev[r->slot], s));
  b->slot_req[r->slot] = r;
  b->slot_state[r->slot] = SLOT_UPLOADING;
  return 0;
}

// ---- work lists ---------------------------------------------------------------

// Weights-pass cost of one iteration of a molecule (the FP64 pass runs at
// about 0.43x the FP32 rate).
double weights_cost(const Request *r) { return (double)r->N * (double)r->F * (r->stage == HP_RUN_FP64 ? 2.3 : 1.0); }

// Launches of one group of molecules; their work entries go to h_work[w...].
// Returns false, with nothing written past work_cap, if the entries do not fit.
bool plan_group(hp_batch *b, const std::vector<int> &active, std::vector<Launch> &plan, int64_t &w) {
  plan.clear();
  auto room = [&](int64_t n) { return w + n <= b->work_cap; };
  // weights launches: one per (precision, function kind, workgroup shape),
  // the FP64 ones first: an FP32-stage molecule with the defect correction
  // is in both, and on its refresh iterations its FP32 pass reads the
  // averages of its FP64 pass
  std::map<std::tuple<int, int, int, int>, std::vector<int>> groups;
  for (int sl : active) {
    const Request *r = b->slot_req[sl];
    const int pt = r->wshape.pt, bs = r->wshape.bs;
    const int fk = r->scheme01 == 0 ? HP_FN_SLATER : HP_FN_GAUSS;
    if (r->stage == HP_RUN_FP64 || r->corr_period > 0) groups[{0, fk, bs, pt}].push_back(sl);
    if (r->stage == HP_RUN_FP32) groups[{1, fk, bs, pt}].push_back(sl);
  }
  for (auto &g : groups) {
    Launch L{};
    L.kind = 0;
    L.fp64 = std::get<0>(g.first) == 0;
    L.fk = std::get<1>(g.first);
    L.bs = std::get<2>(g.first);
    L.pt = std::get<3>(g.first);
    L.off = w;
    int kmax = 1, amax = 1;
    for (int sl : g.second) {
      const Request *r = b->slot_req[sl];
      if (!room((int64_t)r->grp_first.size())) return false;
      for (int s : r->grp_first) b->h_work[w++] = make_int2(r->s0 + s, sl);
      kmax = std::max(kmax, r->maxK);
      amax = std::max(amax, r->A);
    }
    L.len = w - L.off;
    L.km = weights_km(kmax, L.fk);
    L.tile = L.fp64 ? weights_tile_size<double>(L.km, amax) : weights_tile_size<float>(L.km, amax);
    L.lds = hp::weights_lds(L.fp64 ? weights_lds_records<double>(L.km, L.tile, L.fk)
                                   : weights_lds_records<float>(L.km, L.tile, L.fk),
                            hp::WeightsShape{L.pt, L.bs, 1});
    plan.push_back(L);
  }
  // MBIS inner solver: one launch per layout (a molecule's own, as alone);
  // GISA QP: one launch per points-per-lane
  for (int ppl : {2, 3, 4, 1}) {
    Launch L{};
    L.kind = 1;
    L.off = w;
    L.ppl = ppl;
    int kmax = 1, rmax = 1;
    for (int sl : active) {
      const Request *r = b->slot_req[sl];
      if (r->scheme01 != 0 || hp::mbis_ppl(r->maxR) != ppl) continue;
      if (!room(r->A)) return false;
      for (int a = 0; a < r->A; ++a) b->h_work[w++] = make_int2(r->a0 + a, sl);
      kmax = std::max(kmax, r->maxK);
      rmax = std::max(rmax, r->maxR);
    }
    L.len = w - L.off;
    L.km = kmax <= 2 ? 2 : (kmax <= 3 ? 3 : 7);
    L.block = ppl > 1 ? 64 : ((rmax + 63) / 64) * 64;
    if (L.len) plan.push_back(L);
  }
  for (int ppl : {2, 4, 8}) {
    Launch L{};
    L.kind = 2;
    L.ppl = ppl;
    L.off = w;
    for (int sl : active) {
      const Request *r = b->slot_req[sl];
      if (r->scheme01 != 1 || hp::pick_ppl(r->maxR) != ppl) continue;
      if (!room(r->A)) return false;
      for (int a = 0; a < r->A; ++a) b->h_work[w++] = make_int2(r->a0 + a, sl);
    }
    L.len = w - L.off;
    if (L.len) plan.push_back(L);
  }
  return true;
}

// Split the molecules in flight into cfg.groups groups of about equal weights
// cost (largest first, each to the lightest group) and build their launches.
int build_plan(hp_batch *b) {
  std::vector<int> active;
  for (int sl = 0; sl < b->cfg.max_inflight; ++sl)
    if (b->slot_state[sl] == SLOT_ACTIVE) active.push_back(sl);
  std::sort(active.begin(), active.end(), [&](int x, int y) {
    const double cx = weights_cost(b->slot_req[x]), cy = weights_cost(b->slot_req[y]);
    return cx != cy ? cx > cy : x < y;
  });
  const int G = b->cfg.groups;
  std::vector<int> grp[kMaxGroups];
  double load[kMaxGroups] = {};
  for (int sl : active) {
    int g = 0;
    for (int h = 1; h < G; ++h)
      if (load[h] < load[g]) g = h;
    grp[g].push_back(sl);
    load[g] += weights_cost(b->slot_req[sl]);
  }
  int64_t w = 0;
  int ns = 0;
  for (int g = 0; g < kMaxGroups; ++g) {
    std::sort(grp[g].begin(), grp[g].end());
    b->slot_off[g] = ns;
    b->nslot[g] = (int)grp[g].size();
    for (int sl : grp[g]) b->h_slot_list[ns++] = sl;
    if (!plan_group(b, grp[g], b->plan[g], w)) return fail(HP_E_ARG, "work list overflow");
  }
  b->nslot_list = ns;
  if (w) HP_HIP(hipMemcpyAsync(b->d_work, b->h_work, w * sizeof(int2), hipMemcpyHostToDevice, b->cs));
  if (ns) HP_HIP(hipMemcpyAsync(b->d_slot_list, b->h_slot_list, ns * sizeof(int), hipMemcpyHostToDevice, b->cs));
  return 0;
}

template <typename R, int FK, int PT, int KM>
void launch_w(hp_batch *b, const Launch &L, hipStream_t s) {
  hipLaunchKernelGGL((k_weights<R, FK, false, PT, KM>), dim3((unsigned)L.len), dim3(L.bs), L.lds, s, nullptr,
                     b->mol(L.off, 0), b->fn(), b->rhobar, b->ent_part, nullptr, b->cfg.density_cutoff, L.tile,
                     b->corr, HP_CORR_SLOT);
}

template <typename R, int FK, int PT>
void launch_w_pt(hp_batch *b, const Launch &L, hipStream_t s) {
  if (L.km == 2)
    launch_w<R, FK, PT, 2>(b, L, s);
  else if (L.km == 4)
    launch_w<R, FK, PT, 4>(b, L, s);
  else if (L.km == 8)
    launch_w<R, FK, PT, 8>(b, L, s);
  else
    launch_w<R, FK, PT, 16>(b, L, s);
}

template <typename R>
void launch_w_r(hp_batch *b, const Launch &L, hipStream_t s) {
  if (L.fk == HP_FN_SLATER) {
    if (L.pt == 1)
      launch_w_pt<R, HP_FN_SLATER, 1>(b, L, s);
    else
      launch_w_pt<R, HP_FN_SLATER, 2>(b, L, s);
  } else {
    if (L.pt == 1)
      launch_w_pt<R, HP_FN_GAUSS, 1>(b, L, s);
    else
      launch_w_pt<R, HP_FN_GAUSS, 2>(b, L, s);
  }
}

// Event e of group g in tick t of a step: 0 start, 1 after k_batch_begin,
// 2 after the weights launches, 3 after the per-atom launches, 4 end.
inline hipEvent_t &tick_event(hp_batch *b, int t, int g, int e) {
  return b->tev[((size_t)t * kMaxGroups + g) * 5 + e];
}

// One outer iteration of every molecule in flight (tick t of the step).  The
// weights launches of the groups take turns within the tick (group g after
// the previous group of the same tick), so that each group's latency-bound
// per-atom solves run beside the other groups' weights passes.  The first
// group of a tick does not wait for the last group of the previous one: a
// group held up by a long inner solve then delays only the groups after it
// in its own tick.
int enqueue_tick(hp_batch *b, int t) {
  int used[kMaxGroups], nused = 0;
  for (int g = 0; g < kMaxGroups; ++g)
    if (b->nslot[g] > 0) used[nused++] = g;
  for (int u = 0; u < nused; ++u) {
    const int g = used[u];
    const int ns = b->nslot[g];
    hipStream_t s = b->gs[g];
    if (u > 0) HP_HIP(hipStreamWaitEvent(s, tick_event(b, t, used[u - 1], 2), 0));
    const int *slot_list = b->d_slot_list + b->slot_off[g];
    HP_HIP(hipEventRecord(tick_event(b, t, g, 0), s));
    hipLaunchKernelGGL(k_batch_begin, dim3(ns), dim3(256), 0, s, b->slots, slot_list, b->par_m, b->old_m, b->par_l,
                       b->old_l, b->fn_norm, b->fn_alpha, b->coef, b->alpha);
    HP_LAUNCHED();
    HP_HIP(hipEventRecord(tick_event(b, t, g, 1), s));
    bool inner_started = false;
    for (const Launch &L : b->plan[g]) {
      if (L.kind != 0 && !inner_started) {
        HP_HIP(hipEventRecord(tick_event(b, t, g, 2), s));
        inner_started = true;
      }
      if (L.kind == 0) {
        if (L.fp64)
          launch_w_r<double>(b, L, s);
        else
          launch_w_r<float>(b, L, s);
      } else if (L.kind == 1) {
        const DevMol m = b->mol(L.off, (int)L.len);
#define HP_BMBIS(KMV, PPLV)                                                                                         \
  hipLaunchKernelGGL((k_mbis_inner<KMV, PPLV>), dim3((unsigned)L.len), dim3(L.block), 0, s, nullptr, m, b->rhobar,  \
                     b->par_m, b->old_m, b->fn_off, b->cfg.inner_threshold, b->cfg.inner_maxiter, b->inner_niter,     \
                     b->status, b->charges, b->change_part, b->cfg.inner_budget, b->ist_it, b->ist_oldpro)
#define HP_BMBIS_K(KMV) \
  do {                     \
    if (L.ppl == 2)        \
      HP_BMBIS(KMV, 2);    \
    else if (L.ppl == 3)   \
      HP_BMBIS(KMV, 3);    \
    else if (L.ppl == 4)   \
      HP_BMBIS(KMV, 4);    \
    else                   \
      HP_BMBIS(KMV, 1);    \
  } while (0)
        if (L.km == 2)
          HP_BMBIS_K(2);
        else if (L.km == 3)
          HP_BMBIS_K(3);
        else
          HP_BMBIS_K(7);
#undef HP_BMBIS_K
#undef HP_BMBIS
      } else {
        const DevMol m = b->mol(L.off, (int)L.len);
        const dim3 grid((unsigned)((L.len + 3) / 4)), block(256);
#define HP_BQP(PPLV)                                                                                               \
  hipLaunchKernelGGL((k_gisa_qp<PPLV>), grid, block, 0, s, nullptr, m, b->rhobar, b->par_l, b->old_l, b->fn_off,   \
                     b->fn_norm, b->fn_alpha, b->fn_order, b->qpP, b->qp_off, b->cfg.population_cutoff,           \
                     b->cfg.negative_cutoff, b->inner_niter, b->status, b->charges, b->change_part)
        if (L.ppl == 2)
          HP_BQP(2);
        else if (L.ppl == 4)
          HP_BQP(4);
        else
          HP_BQP(8);
#undef HP_BQP
      }
      HP_LAUNCHED();
    }
    if (!inner_started) HP_HIP(hipEventRecord(tick_event(b, t, g, 2), s));
    HP_HIP(hipEventRecord(tick_event(b, t, g, 3), s));
    hipLaunchKernelGGL(k_batch_finalize, dim3(ns), dim3(256), 0, s, b->slots, slot_list, b->ent_part, b->change_part,
                       b->status, b->status_all, b->par_m, b->old_m, b->par_l, b->old_l, b->ist_it);
    HP_LAUNCHED();
    HP_HIP(hipEventRecord(tick_event(b, t, g, 4), s));
  }
  return 0;
}

// GPU time of the previous step (after the streams were synchronised).
int account_ticks(hp_batch *b) {
  float ms = 0.f;
  HP_HIP(hipEventElapsedTime(&ms, b->ev_step[0], b->ev_step[1]));
  b->stats.tick_ms += ms;
  b->stats.ticks += b->cfg.ticks_per_step;
  if (b->cfg.use_graphs) return 0;  // events inside a replayed graph are not timed
  for (int t = 0; t < b->cfg.ticks_per_step; ++t) {
    for (int g = 0; g < kMaxGroups; ++g) {
      if (b->nslot[g] == 0) continue;
      float ms_w = 0.f, ms_i = 0.f;
      HP_HIP(hipEventElapsedTime(&ms_w, tick_event(b, t, g, 1), tick_event(b, t, g, 2)));
      HP_HIP(hipEventElapsedTime(&ms_i, tick_event(b, t, g, 2), tick_event(b, t, g, 3)));
      b->stats.weights_ms += ms_w;
      b->stats.inner_ms += ms_i;
    }
  }
  return 0;
}

// The ticks of one step and the read-back of the slot records; the other
// groups' streams fork from and join the compute stream.
int enqueue_step_work(hp_batch *b) {
  const int G = b->cfg.groups;
  HP_HIP(hipEventRecord(b->ev_fork, b->cs));
  for (int g = 1; g < G; ++g) HP_HIP(hipStreamWaitEvent(b->gs[g], b->ev_fork, 0));
  for (int t = 0; t < b->cfg.ticks_per_step; ++t) {
    int rc = enqueue_tick(b, t);
    if (rc) return rc;
  }
  for (int g = 1; g < G; ++g) {
    HP_HIP(hipEventRecord(b->ev_join[g], b->gs[g]));
    HP_HIP(hipStreamWaitEvent(b->cs, b->ev_join[g], 0));
  }
  HP_HIP(hipMemcpyAsync(b->h_slots, b->slots, b->cfg.max_inflight * sizeof(HpSlot), hipMemcpyDeviceToHost, b->cs));
  return 0;
}

int drop_graph(hp_batch *b) {
  if (b->gexec) HP_HIP(hipGraphExecDestroy(b->gexec));
  if (b->graph) HP_HIP(hipGraphDestroy(b->graph));
  b->gexec = nullptr;
  b->graph = nullptr;
  return 0;
}

int run_step_work(hp_batch *b) {
  HP_HIP(hipEventRecord(b->ev_step[0], b->cs));
  if (!b->cfg.use_graphs) {
    int rc = enqueue_step_work(b);
    if (rc) return rc;
  } else {
    if (!b->gexec) {
      HP_HIP(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_step_work(b);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->cs, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(HP_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      b->graph = g;
      HP_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    }
    HP_HIP(hipGraphLaunch(b->gexec, b->cs));
  }
  HP_HIP(hipEventRecord(b->ev_step[1], b->cs));
  return 0;
}

// Copy a stopped molecule's results out and free its ranges.
int retire(hp_batch *b, int sl) {
  Request *r = b->slot_req[sl];
  const HpSlot &h = b->h_slots[sl];
  const double *par = r->scheme01 == 0 ? b->par_m + 2 * (int64_t)r->f0 : b->par_l + r->f0;
  HP_HIP(hipMemcpyAsync(r->propars_out, par, r->npar * sizeof(double), hipMemcpyDeviceToHost, b->cs));
  HP_HIP(hipMemcpyAsync(r->charges_out, b->charges + r->a0, r->A * sizeof(double), hipMemcpyDeviceToHost, b->cs));
  hp_batch_result res{};
  res.ticket = r->ticket;
  res.event = h.event == HP_EV_CONVERGED ? HP_BATCH_CONVERGED : (h.event == HP_EV_MAXITER ? HP_BATCH_MAXITER : HP_BATCH_ERROR);
  res.niter = h.niter;
  res.status = h.any;
  res.status_last = h.err;
  res.change = h.change;
  res.entropy = h.entropy;
  b->done.push_back(res);
  b->stats.completed += 1;
  b->slot_state[sl] = SLOT_FREE;
  b->slot_req[sl] = nullptr;
  unplace(b, r);
  delete r;
  return 0;
}

}  // namespace

extern "C" {

int hp_batch_create(int device, const hp_batch_config *cfg, hp_batch **out) {
  hp::g_err.clear();
  if (!cfg || !out) return fail(HP_E_ARG, "NULL argument");
  *out = nullptr;
  if (cfg->point_capacity <= 0 || cfg->atom_capacity <= 1 || cfg->shell_capacity <= 1 || cfg->function_capacity <= 0 ||
      cfg->max_inflight <= 0 || cfg->qp_capacity < 0 || cfg->max_molecule_points <= 0 || cfg->ticks_per_step <= 0)
    return fail(HP_E_ARG, "batch capacities must be positive");
  if (cfg->max_inflight > 4096) return fail(HP_E_ARG, "at most 4096 molecules in flight");
  if (cfg->groups < 0 || cfg->groups > kMaxGroups) return fail(HP_E_ARG, "groups must be in [0, 4]");
  if (cfg->inner_budget < 0) return fail(HP_E_ARG, "inner_budget must be >= 0");
  hp_batch *b = new hp_batch();
  b->device = device;
  b->cfg = *cfg;
  if (b->cfg.groups == 0) b->cfg.groups = default_groups();
  b->cfg.max_molecule_points = std::min(cfg->max_molecule_points, cfg->point_capacity);
  auto bail = [&](int rc) {
    free_batch(b);
    return rc;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return bail(fail(HP_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
  if ((e = hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&b->us, hipStreamNonBlocking)) != hipSuccess)
    return bail(fail(HP_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  b->gs[0] = b->cs;
  for (int g = 1; g < b->cfg.groups; ++g)
    if ((e = hipStreamCreateWithFlags(&b->gs[g], hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&b->ev_join[g])) != hipSuccess)
      return bail(fail(HP_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  const int64_t NP = cfg->point_capacity, NA = cfg->atom_capacity, NS = cfg->shell_capacity,
                NF = cfg->function_capacity, NQ = std::max<int64_t>(1, cfg->qp_capacity), MI = cfg->max_inflight,
                MP = b->cfg.max_molecule_points;
  // work entries: per molecule its weights workgroups (at most one per radial
  // shell, twice for an FP32-stage molecule with the FP64 correction, which
  // is in both weights launches) and one per atom for its solver
  b->work_cap = 2 * NS + NA;
  int rc = 0;
  if ((rc = dalloc(&b->px, NP)) || (rc = dalloc(&b->py, NP)) || (rc = dalloc(&b->pz, NP)) ||
      (rc = dalloc(&b->rho, NP)) || (rc = dalloc(&b->wmol, NP)) || (rc = dalloc(&b->wsph, NP)) ||
      (rc = dalloc(&b->shell_atom, NS)) || (rc = dalloc(&b->shell_off, NS)) || (rc = dalloc(&b->shell_r, NS)) ||
      (rc = dalloc(&b->shell_w, NS)) || (rc = dalloc(&b->rhobar, NS)) || (rc = dalloc(&b->ent_part, NS)) ||
      (rc = dalloc(&b->corr, NS)) ||
      (rc = dalloc(&b->cx, NA)) || (rc = dalloc(&b->cy, NA)) || (rc = dalloc(&b->cz, NA)) ||
      (rc = dalloc(&b->pseudo, NA)) || (rc = dalloc(&b->charges, NA)) || (rc = dalloc(&b->change_part, NA)) ||
      (rc = dalloc(&b->atom_shell_off, NA)) || (rc = dalloc(&b->fn_off, NA)) || (rc = dalloc(&b->inner_niter, NA)) ||
      (rc = dalloc(&b->status, NA)) || (rc = dalloc(&b->status_all, NA)) || (rc = dalloc(&b->qp_off, NA)) ||
      (rc = dalloc(&b->coef, NF)) || (rc = dalloc(&b->alpha, NF)) || (rc = dalloc(&b->fn_norm, NF)) ||
      (rc = dalloc(&b->fn_alpha, NF)) || (rc = dalloc(&b->fn_order, NF)) || (rc = dalloc(&b->par_l, NF)) ||
      (rc = dalloc(&b->old_l, NF)) || (rc = dalloc(&b->par_m, 2 * NF)) || (rc = dalloc(&b->old_m, 2 * NF)) ||
      (rc = dalloc(&b->qpP, NQ)) || (rc = dalloc(&b->slots, MI)) || (rc = dalloc(&b->st_points, 3 * MP)) ||
      (rc = dalloc(&b->st_coords, 3 * (NA - 1))) || (rc = dalloc(&b->st_shell_atom, NS)) ||
      (rc = dalloc(&b->st_fn_off, NA)) || (rc = dalloc(&b->st_shell_off, NS)) ||
      (rc = dalloc(&b->d_work, b->work_cap)) || (rc = dalloc(&b->d_slot_list, MI)) ||
      (rc = dalloc(&b->evals_done, HP_EVAL_SLOTS)) || (rc = dalloc(&b->ist_it, NA)) ||
      (rc = dalloc(&b->ist_oldpro, NS)))
    return bail(rc);
  if ((e = hipMemset(b->evals_done, 0, HP_EVAL_SLOTS * sizeof(unsigned long long))) != hipSuccess)
    return bail(fail(HP_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e)));
  if ((e = hipHostMalloc((void **)&b->h_slots, MI * sizeof(HpSlot), hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc((void **)&b->h_slot_init, MI * sizeof(HpSlot), hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc((void **)&b->h_work, b->work_cap * sizeof(int2), hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc((void **)&b->h_slot_list, MI * sizeof(int), hipHostMallocDefault)) != hipSuccess)
    return bail(fail(HP_E_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e)));
  std::memset(b->h_slots, 0, MI * sizeof(HpSlot));
  b->slot_state.assign(MI, SLOT_FREE);
  b->slot_req.assign(MI, nullptr);
  b->up_ev.assign(MI, nullptr);
  for (auto &ev : b->up_ev)
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
      return bail(fail(HP_E_HIP, "hipEventCreate failed"));
  b->tev.assign(5 * (size_t)kMaxGroups * cfg->ticks_per_step, nullptr);
  for (auto &ev : b->tev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(fail(HP_E_HIP, "hipEventCreate failed"));
  for (hipEvent_t *ev : {&b->ev_fork, &b->ev_step[0], &b->ev_step[1]})
    if ((e = hipEventCreate(ev)) != hipSuccess) return bail(fail(HP_E_HIP, "hipEventCreate failed"));
  b->last_niter.assign(MI, 0);
  for (int i = (int)MI - 1; i >= 0; --i) b->free_slots.push_back(i);
  b->pts.init(NP);
  b->atoms.init(NA);
  b->shells.init(NS);
  b->fns.init(NF);
  b->qps.init(NQ);
  // every slot starts idle; fill kernels only (no memset nodes in graphs)
  e = hipMemcpyAsync(b->slots, b->h_slots, MI * sizeof(HpSlot), hipMemcpyHostToDevice, b->cs);
  if (e == hipSuccess) e = hipStreamSynchronize(b->cs);
  if (e != hipSuccess) return bail(fail(HP_E_HIP, std::string("initialisation: ") + hipGetErrorString(e)));
  *out = b;
  return 0;
}

int hp_batch_destroy(hp_batch *b) {
  if (!b) return fail(HP_E_ARG, "NULL batch");
  (void)hipSetDevice(b->device);
  for (int g = 0; g < kMaxGroups; ++g)
    if (b->gs[g]) (void)hipStreamSynchronize(b->gs[g]);
  (void)hipStreamSynchronize(b->us);
  free_batch(b);
  return 0;
}

int hp_batch_submit(hp_batch *b, const hp_molecule_desc *md, const hp_scheme_desc *sd, const double *propars0,
                    int precision, int n_polish, double *propars_out, double *charges_out, int64_t *ticket) {
  hp::g_err.clear();
  if (!b || !md || !sd || !ticket) return fail(HP_E_ARG, "NULL argument");
  Request *r = new Request();
  r->md = *md;
  r->sd = *sd;
  r->propars0 = propars0;
  r->propars_out = propars_out;
  r->charges_out = charges_out;
  r->precision = precision;
  r->n_polish = n_polish;
  int rc = inspect(b, r);
  if (rc) {
    delete r;
    return rc;
  }
  r->ticket = b->next_ticket++;
  *ticket = r->ticket;
  b->queue.push_back(r);
  return 0;
}

int hp_batch_step(hp_batch *b, int32_t *n_inflight, int32_t *n_queued) {
  if (!b) return fail(HP_E_ARG, "NULL batch");
  HP_HIP(hipSetDevice(b->device));
  const int MI = b->cfg.max_inflight;
  // 1. the previous ticks and their slot read-back
  HP_HIP(hipStreamSynchronize(b->cs));
  // 2. retire, and collect stage switches
  bool changed = false;
  int nretired = 0;
  if (b->ticks_in_flight) {
    int rc = account_ticks(b);
    if (rc) return rc;
    for (int sl = 0; sl < MI; ++sl) {
      if (b->slot_state[sl] != SLOT_ACTIVE) continue;
      const HpSlot &h = b->h_slots[sl];
      {
        // the iterations of the previous step ran in the stage the host knew
        const Request *r = b->slot_req[sl];
        const int d = h.niter - b->last_niter[sl];
        b->last_niter[sl] = h.niter;
        b->stats.iterations += d;
        const double pts = (double)d * (double)r->N, ev = pts * (double)r->F;
        if (r->stage == HP_RUN_FP32) {
          b->stats.points_fp32 += pts;
          b->stats.evals_fp32 += ev;
          if (r->corr_period > 0 && d > 0) {
            // refresh iterations among [niter - d, niter) also ran in FP64
            const int m = r->corr_period, n1 = h.niter, n0 = n1 - d;
            const int nr = (n1 + m - 1) / m - (n0 + m - 1) / m;
            b->stats.points_fp64 += (double)nr * r->N;
            b->stats.evals_fp64 += (double)nr * r->N * r->F;
          }
        } else {
          b->stats.points_fp64 += pts;
          b->stats.evals_fp64 += ev;
        }
      }
      if (h.run != HP_RUN_IDLE || h.event == HP_EV_RUNNING) continue;
      if (h.event == HP_EV_SWITCH) {
        b->slot_req[sl]->stage = HP_RUN_FP64;
        b->pending_fp64.push_back(sl);
      } else {
        int rc = retire(b, sl);
        if (rc) return rc;
        ++nretired;
      }
      changed = true;
    }
  }
  // the results are out before their ranges can be refilled or reported
  if (nretired) HP_HIP(hipStreamSynchronize(b->cs));
  // 3. molecules whose upload finished join (with nothing to compute, wait
  // for the uploads rather than return at once)
  bool any_active = false, any_uploading = false;
  for (int sl = 0; sl < MI; ++sl) {
    any_active = any_active || b->slot_state[sl] == SLOT_ACTIVE;
    any_uploading = any_uploading || b->slot_state[sl] == SLOT_UPLOADING;
  }
  if (!any_active && any_uploading) HP_HIP(hipStreamSynchronize(b->us));
  std::vector<int> joining;
  for (int sl = 0; sl < MI; ++sl) {
    if (b->slot_state[sl] != SLOT_UPLOADING) continue;
    const hipError_t q = hipEventQuery(b->up_ev[sl]);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return fail(HP_E_HIP, std::string("upload: ") + hipGetErrorString(q));
    b->slot_state[sl] = SLOT_ACTIVE;
    b->last_niter[sl] = 0;
    joining.push_back(sl);
    changed = true;
  }
  // 4. work lists and stage changes
  if (changed) {
    int rc = drop_graph(b);
    if (rc) return rc;
    rc = build_plan(b);
    if (rc) return rc;
    // the joiners' slot records were written by the upload stream
    auto set_run = [&](std::vector<int> &ids, int run) -> int {
      for (size_t i = 0; i < ids.size(); i += 62) {
        HpRunSet rs{};
        rs.run = run;
        rs.n = (int)std::min<size_t>(62, ids.size() - i);
        for (int k = 0; k < rs.n; ++k) rs.slot[k] = ids[i + k];
        hipLaunchKernelGGL(k_batch_set_run, dim3(1), dim3(64), 0, b->cs, b->slots, rs);
        HP_LAUNCHED();
      }
      return 0;
    };
    std::vector<int> j32, j64;
    for (int sl : joining) (b->slot_req[sl]->stage == HP_RUN_FP32 ? j32 : j64).push_back(sl);
    if ((rc = set_run(j32, HP_RUN_FP32)) || (rc = set_run(j64, HP_RUN_FP64)) || (rc = set_run(b->pending_fp64, HP_RUN_FP64)))
      return rc;
    b->pending_fp64.clear();
  }
  // 5. the next ticks
  b->ticks_in_flight = b->nslot_list > 0;
  if (b->ticks_in_flight) {
    int rc = run_step_work(b);
    if (rc) return rc;
  }
  // 6. admissions (upload stream, overlapping the ticks): freed ranges are
  // read by none of the enqueued kernels and their results are already out
  while (!b->queue.empty()) {
    Request *r = b->queue.front();
    if (!place(b, r)) break;
    b->queue.pop_front();
    int rc = upload(b, r);
    if (rc) return rc;
  }
  if (n_inflight) {
    int n = 0;
    for (int sl = 0; sl < MI; ++sl) n += b->slot_state[sl] != SLOT_FREE;
    *n_inflight = n;
  }
  if (n_queued) *n_queued = (int32_t)b->queue.size();
  return 0;
}

int hp_host_alloc(int64_t bytes, void **ptr) {
  if (!ptr || bytes <= 0) return fail(HP_E_ARG, "bad host allocation request");
  *ptr = nullptr;
  HP_HIP(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault));
  return 0;
}

int hp_host_free(void *ptr) {
  if (ptr) HP_HIP(hipHostFree(ptr));
  return 0;
}

int hp_batch_get_stats(hp_batch *b, hp_batch_stats *out) {
  if (!b || !out) return fail(HP_E_ARG, "NULL argument");
  HP_HIP(hipSetDevice(b->device));
  for (int g = 0; g < kMaxGroups; ++g)
    if (b->gs[g]) HP_HIP(hipStreamSynchronize(b->gs[g]));
  double d32 = 0.0, d64 = 0.0;
  if (int rc = hp::read_eval_counters(b->evals_done, &d32, &d64)) return rc;
  *out = b->stats;
  out->done_fp32 = d32;
  out->done_fp64 = d64;
  return 0;
}

int hp_batch_poll(hp_batch *b, hp_batch_result *out, int32_t max, int32_t *n) {
  if (!b || !n || (max > 0 && !out)) return fail(HP_E_ARG, "NULL argument");
  int k = 0;
  while (k < max && !b->done.empty()) {
    out[k++] = b->done.front();
    b->done.pop_front();
  }
  *n = k;
  return 0;
}

}  // extern "C"
