// Tabulated pro-atoms: ISA (isa.py:37-122), Hirshfeld-I (hirshfeld_i.py:36-176)
// and Hirshfeld (hirshfeld.py:127-193).
//
// Every atom's pro-atom is a piecewise cubic over its radial knots (the
// atom's radial grid for ISA, its element's pro-atom database grid for
// Hirshfeld and Hirshfeld-I), evaluated at each molecular point by
// proatom_value<.., HP_FN_TABLE> (an interval search in the knots and one
// cubic).  One outer iteration is
//   k_isa_begin / k_hi_begin   build the tables from the parameters
//   k_table_weights            promolecule and AIM weights per point; per
//                              radial shell the spherical average, the
//                              atom-grid population and the entropy partial
//   k_isa_update / k_hi_update the new parameters, charges and change
//   k_finalize                 (shared with the other schemes)
// Every kernel returns at once while the device stop flag is set, so
// iterations are enqueued ahead of the host's decision.
#pragma once
#include "hp_device.h"

#define HP_TAB_MAXN 512  // knots per atom (LDS arrays of the ISA spline build)

// Pro-atom database records of the atoms (Hirshfeld-I): records of atom a are
// [rec_off[a], rec_off[a+1]); record j has the integer charge rec_charge[j],
// density values rec_rho[rec_val[j] + i] and cubic pieces
// rec_c[4 (rec_val[j] + i) ..] on the knots of its atom.
struct DevRec {
  const int *rec_off;
  const int *rec_charge;
  const int64_t *rec_val;
  const double *rec_rho;
  const double *rec_c;
};

// Sum of three values over a workgroup (blockDim multiple of 64); thread 0
// gets them.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *lds) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
    c += __shfl_down(c, o, 64);
  }
  if (lane == 0) {
    lds[3 * wv] = a;
    lds[3 * wv + 1] = b;
    lds[3 * wv + 2] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0, sc = 0.0;
    for (int w = 0; w < nw; ++w) {
      sa += lds[3 * w];
      sb += lds[3 * w + 1];
      sc += lds[3 * w + 2];
    }
    a = sa;
    b = sb;
    c = sc;
  }
}

// ISA: each atom's table is the not-a-knot spline through its parameters
// (the spherical averages at its radial knots, isa.py:65-91 with
// core/stockholder.py:227-264: CubicSpline(rgrid.points, rho)); negative
// values are taken as zero (fix_proatom_rho, core/stockholder.py:201-225).
// One workgroup of 64 per atom: the knots and values are staged in LDS and
// one lane runs the elimination.
static __global__ __launch_bounds__(64) void k_isa_begin(const int *__restrict__ done, const int *__restrict__ tab_off,
                                                          const double *__restrict__ tab_x,
                                                          const double *__restrict__ par, double *__restrict__ old_par,
                                                          double *__restrict__ tab_c) {
  if (done && *done) return;
  __shared__ double sx[HP_TAB_MAXN], sy[HP_TAB_MAXN], scp[HP_TAB_MAXN], srp[HP_TAB_MAXN];
  const int a = blockIdx.x;
  const int k0 = tab_off[a], n = tab_off[a + 1] - k0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = par[k0 + i];
    old_par[k0 + i] = v;
    sx[i] = tab_x[k0 + i];
    sy[i] = v < 0.0 ? 0.0 : v;
  }
  __syncthreads();
  if (threadIdx.x == 0) notaknot_spline(n, sx, sy, scp, srp, tab_c + 4 * (int64_t)k0);
}

// Records and coefficients of the Hirshfeld-I pro-atom of an atom with charge
// q (get_interpolation_info and get_proatom_rho, hirshfeld_i.py:116-134;
// eval_proatom 148-158): with i = floor(q), x = q - i and the pseudo
// population P = Z_pseudo - i, the pro-atom is (1 - x) rec(i) when P == 1 or
// x == 0, else (1 - x) rec(i) + x rec(i + 1) when P > 1.  Returns false when
// the reference raises: P <= 0 (ValueError), 0 < P < 1 (no pro-atom), or a
// record that the database lacks (KeyError).
__device__ __forceinline__ bool hi_records(const DevRec &rec, int a, double pseudo, double q, int &j0, double &w0,
                                           int &j1, double &w1) {
  if (!(q == q) || q > 1.0e9 || q < -1.0e9) return false;
  const double fq = floor(q);
  const int ic = (int)fq;
  const double x = q - fq;
  const double pp = pseudo - fq;
  j0 = -1;
  j1 = -1;
  w0 = 1.0 - x;
  w1 = 0.0;
  if (!(pp > 0.0)) return false;
  for (int j = rec.rec_off[a]; j < rec.rec_off[a + 1]; ++j) {
    if (rec.rec_charge[j] == ic) j0 = j;
    if (rec.rec_charge[j] == ic + 1) j1 = j;
  }
  if (j0 < 0) return false;
  if (pp == 1.0 || x == 0.0) {
    j1 = -1;
    return true;
  }
  if (pp > 1.0) {
    w1 = x;
    return j1 >= 0;
  }
  return false;
}

// Hirshfeld-I: the tables and the knot densities (for the change) of the
// pro-atoms at the current charges.  One workgroup of 64 per atom.
static __global__ __launch_bounds__(64) void k_hi_begin(const int *__restrict__ done, int natom,
                                                         const int *__restrict__ tab_off, DevRec rec,
                                                         const double *__restrict__ pseudo,
                                                         const double *__restrict__ par, double *__restrict__ old_par,
                                                         double *__restrict__ tab_c, double *__restrict__ cur_rho,
                                                         int *__restrict__ status) {
  if (done && *done) return;
  const int a = blockIdx.x;
  const int k0 = tab_off[a], n = tab_off[a + 1] - k0;
  const double q = par[a];
  if (threadIdx.x == 0) {
    old_par[a] = q;
    status[a] = 0;
  }
  int j0, j1;
  double w0, w1;
  const bool ok = hi_records(rec, a, pseudo[a], q, j0, w0, j1, w1);
  if (!ok && threadIdx.x == 0) status[a] = HP_ST_NORECORD;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double rho = 0.0, c[4] = {0.0, 0.0, 0.0, 0.0};
    if (ok) {
      const int64_t v0 = rec.rec_val[j0] + i;
      rho = w0 * rec.rec_rho[v0];
#pragma unroll
      for (int k = 0; k < 4; ++k) c[k] = w0 * rec.rec_c[4 * v0 + k];
      if (j1 >= 0) {
        const int64_t v1 = rec.rec_val[j1] + i;
        rho += w1 * rec.rec_rho[v1];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] += w1 * rec.rec_c[4 * v1 + k];
      }
    }
    cur_rho[k0 + i] = rho;
#pragma unroll
    for (int k = 0; k < 4; ++k) tab_c[4 * (int64_t)(k0 + i) + k] = c[k];
  }
}

// update_at_weights (core/stockholder.py:352-384) with tabulated pro-atoms,
// one workgroup per radial shell: per point the promolecule (atom order,
// 1e-100 in every pro-atom and after every atom) and the shell atom's weight
// clip(rho0_a / rho0, 0, 1); per shell the spherical average of w rho
// (sum_j wsph_j w_j rho_j), the atom-grid population sum_j wat_j w_j rho_j and
// the entropy partial (stockholder.py:145-151).  promol (may be null) keeps
// the promolecule per point.
static __global__ __launch_bounds__(256) void k_table_weights(const int *__restrict__ done, DevMol m, DevFn fn, double dc,
                                                               double *__restrict__ rhobar, double *__restrict__ pop_part,
                                                               double *__restrict__ ent_part,
                                                               double *__restrict__ promol_out) {
  if (done && *done) return;
  __shared__ double lds[48];
  const int s = blockIdx.x;
  const int a = m.shell_atom[s];
  const int64_t p0 = m.shell_off[s], p1 = m.shell_off[s + 1];
  double sph = 0.0, pop = 0.0, ent = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    double promol, own;
    promol_at<double, HP_FN_TABLE>(m, fn, a, m.natom, nullptr, m.px[p], m.py[p], m.pz[p], promol, own);
    if (promol_out) promol_out[p] = promol;
    const double rho = m.rho[p];
    const double f = aim_weight(own, promol) * rho;
    sph += m.wsph[p] * f;
    pop += m.wat[p] * f;
    const bool sick = (promol < dc) || (rho < dc);
    ent += sick ? 0.0 : m.wmol[p] * rho * log(rho / promol);
  }
  block_sum3(sph, pop, ent, lds);
  if (threadIdx.x == 0) {
    rhobar[s] = sph;
    pop_part[s] = pop;
    ent_part[s] = ent;
  }
}

// ISA update (isa.py:102-122), one wave per atom: the parameters become the
// spherical averages clipped below at 1e-100; the charge is
// Z_pseudo - sum_i w_i 4 pi r_i^2 rhobar_i and the change partial
// sum_i w_i 4 pi r_i^2 (new_i - old_i)^2 (iterstock.py:32-45) over the knots,
// which are the atom's radial shells.
static __global__ __launch_bounds__(64) void k_isa_update(const int *__restrict__ done, DevMol m,
                                                           const int *__restrict__ tab_off,
                                                           const double *__restrict__ tab_x,
                                                           const double *__restrict__ tab_w,
                                                           const double *__restrict__ rhobar, double *__restrict__ par,
                                                           const double *__restrict__ old_par,
                                                           double *__restrict__ charges,
                                                           double *__restrict__ change_part, int *__restrict__ status) {
  if (done && *done) return;
  const int a = blockIdx.x;
  const int k0 = tab_off[a], n = tab_off[a + 1] - k0;
  const int s0 = m.atom_shell_off[a];
  double pop = 0.0, msd = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) {
    const double r = tab_x[k0 + i];
    const double v = fmax(rhobar[s0 + i], 1e-100);
    const double w4 = tab_w[k0 + i] * (4.0 * HP_PI * r * r);
    par[k0 + i] = v;
    pop += w4 * v;
    const double d = v - old_par[k0 + i];
    msd += w4 * d * d;
  }
  pop = wave_sum(pop);
  msd = wave_sum(msd);
  if (threadIdx.x == 0) {
    charges[a] = m.pseudo[a] - pop;
    change_part[a] = msd;
    status[a] = 0;
  }
}

// Hirshfeld-I update (hirshfeld_i.py:166-172 with core/base.py:259-264), one
// wave per atom: the pseudo population sum over the atom's shells of the
// atom-grid populations, the new charge (which is the parameter), and the
// change sum_i w_i 4 pi r_i^2 (rho_new_i - rho_old_i)^2 on the record knots.
static __global__ __launch_bounds__(64) void k_hi_update(const int *__restrict__ done, DevMol m,
                                                          const int *__restrict__ tab_off,
                                                          const double *__restrict__ tab_x,
                                                          const double *__restrict__ tab_w, DevRec rec,
                                                          const double *__restrict__ pop_part,
                                                          const double *__restrict__ cur_rho, double *__restrict__ par,
                                                          double *__restrict__ charges, double *__restrict__ change_part,
                                                          int *__restrict__ status) {
  if (done && *done) return;
  const int a = blockIdx.x;
  double pop = 0.0;
  for (int s = m.atom_shell_off[a] + threadIdx.x; s < m.atom_shell_off[a + 1]; s += 64) pop += pop_part[s];
  pop = wave_sum(pop);
  const double q = m.pseudo[a] - pop;
  int j0, j1;
  double w0, w1;
  const bool ok = hi_records(rec, a, m.pseudo[a], q, j0, w0, j1, w1);
  const int k0 = tab_off[a], n = tab_off[a + 1] - k0;
  double msd = 0.0;
  if (ok) {
    for (int i = threadIdx.x; i < n; i += 64) {
      const double r = tab_x[k0 + i];
      double rho = w0 * rec.rec_rho[rec.rec_val[j0] + i];
      if (j1 >= 0) rho += w1 * rec.rec_rho[rec.rec_val[j1] + i];
      const double d = rho - cur_rho[k0 + i];
      msd += tab_w[k0 + i] * (4.0 * HP_PI * r * r) * d * d;
    }
  }
  msd = wave_sum(msd);
  if (threadIdx.x == 0) {
    par[a] = q;
    charges[a] = q;
    change_part[a] = msd;
    status[a] |= ok ? 0 : HP_ST_NORECORD;
  }
}
