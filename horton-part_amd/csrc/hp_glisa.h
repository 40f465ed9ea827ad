// gLISA: the global problem over every pro-atom coefficient on the molecular
// grid (glisa.py:248-281, 346-348, 411-479, 617-879).
//
// With M = sum_a K_a basis functions g_i (unit coefficient) and coefficients
// c, the promolecule is rho0 = sum_i c_i g_i (no floors) and, over the points
// where neither rho nor rho0 is below the density cut-off,
//   f      = sum_p w rho ln(rho / rho0)
//   grad_i = -sum_p w rho g_i / rho0
//   H_ij   = sum_p w rho g_i g_j / rho0^2
// k_glisa_points evaluates the basis once per point into a point-major table
// G (row p holds g_0(p) .. g_{M-1}(p), padded to a multiple of 16 with
// zeros), the weights D = w rho / rho0^2 and E = w rho / rho0, and the
// objective partials; k_glisa_grad reduces E g_i per point chunk;
// k_glisa_hess accumulates 16 x 16 tiles of G^T diag(D) G on the matrix cores
// (mfma_f64_16x16x4f64, one wavefront per tile pair and point chunk); fixed-
// order reductions over the chunks follow, so results do not depend on
// scheduling.
#pragma once
#include "hp_device.h"

typedef double hp_gd4 __attribute__((ext_vector_type(4)));

#define HP_GLISA_CHUNK 4096   // points per workgroup of the reductions
#define HP_GLISA_HCHUNK 1024  // points per workgroup of the Hessian passes

// Basis value of function f at distance r (and r^2) from its atom.
template <int FK>
__device__ __forceinline__ double glisa_basis(const DevFn &fn, const double *norm, int f, double r, double r2) {
  double arg;
  if (FK == HP_FN_SLATER) {
    arg = r;
  } else if (FK == HP_FN_GAUSS) {
    arg = r2;
  } else {
    const double n = fn.order[f];
    arg = (n == 1.0) ? r : ((n == 2.0) ? r2 : pow(r, n));
  }
  return norm[f] * exp(-fn.alpha[f] * arg);
}

// One thread per point: G row, D, E and the objective's block partial.  The
// workgroup's 256 rows of G pass through LDS 16 columns at a time, so that G
// is written in whole 128-byte runs of each row (columns M .. Mp are zero).
template <int FK>
__global__ __launch_bounds__(256) void k_glisa_points(DevMol m, DevFn fn, const double *__restrict__ norm,
                                                      const double *__restrict__ coef, int M, int Mp, double dc,
                                                      double *__restrict__ G, double *__restrict__ D,
                                                      double *__restrict__ E, double *__restrict__ obj_part) {
  constexpr int LD = 17;  // LDS row stride (doubles) of the column blocks
  __shared__ double lds[32];
  __shared__ double s_min[4];
  __shared__ double s_t[256 * LD];
  const int64_t pb = (int64_t)blockIdx.x * blockDim.x;
  const int64_t p = pb + threadIdx.x;
  const bool ok = p < m.npoint;
  const int64_t pc = ok ? p : m.npoint - 1;  // idle threads follow the last point (no output)
  const int nrow = (int)(m.npoint - pb < (int64_t)blockDim.x ? m.npoint - pb : (int64_t)blockDim.x);
  // columns [16 blk, 16 blk + 16) of the workgroup's rows, from LDS to G
  auto flush = [&](int blk) {
    __syncthreads();
    for (int q = threadIdx.x; q < 256 * 8; q += blockDim.x) {
      const int row = q >> 3, pr = q & 7;
      if (row < nrow) {
        double2 v;
        v.x = s_t[row * LD + 2 * pr];
        v.y = s_t[row * LD + 2 * pr + 1];
        *reinterpret_cast<double2 *>(G + (pb + row) * Mp + 16 * blk + 2 * pr) = v;
      }
    }
    __syncthreads();
  };
  const double x = m.px[pc], y = m.py[pc], z = m.pz[pc];
  double rho0 = 0.0;
  for (int a = 0; a < m.natom; ++a) {
    const double dx = x - m.cx[a], dy = y - m.cy[a], dz = z - m.cz[a];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    for (int f = fn.fn_off[a]; f < fn.fn_off[a + 1]; ++f) {
      const double g = glisa_basis<FK>(fn, norm, f, r, r2);
      s_t[threadIdx.x * LD + (f & 15)] = g;
      rho0 += coef[f] * g;
      if ((f & 15) == 15) flush(f >> 4);
    }
  }
  if (M & 15) {
    for (int c = M & 15; c < 16; ++c) s_t[threadIdx.x * LD + c] = 0.0;
    flush(M >> 4);
  }
  double fo = 0.0, dummy = 0.0, mn = 1.0e308;
  if (ok) {
    const double rho = m.rho[p], w = m.wmol[p];
    const bool sick = (rho < dc) || (rho0 < dc);
    D[p] = sick ? 0.0 : w * rho / (rho0 * rho0);
    E[p] = sick ? 0.0 : w * rho / rho0;
    fo = sick ? 0.0 : w * rho * log(rho / rho0);
    mn = rho0;
  }
  const double wmin = wave_min(mn);
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = wmin;
  block_sum2(fo, dummy, lds);
  if (threadIdx.x == 0) {
    obj_part[2 * blockIdx.x] = fo;
    obj_part[2 * blockIdx.x + 1] = fmin(fmin(s_min[0], s_min[1]), fmin(s_min[2], s_min[3]));
  }
}

// Gradient partials: workgroup c sums E g_i over its chunk for every i.
// With Mp <= 256 the workgroup is ng = 256 / Mp groups of Mp threads; thread
// i of group h adds E(p) g_i(p) over the points p0 + h, p0 + h + ng, ... (row
// p of G is read by a group as one contiguous run), and the groups' sums are
// added in group order.  Wider tables take columns i, i + 256, ... per thread.
__global__ __launch_bounds__(256) void k_glisa_grad(int64_t npoint, int M, int Mp, const double *__restrict__ G,
                                                    const double *__restrict__ E, double *__restrict__ grad_part) {
  __shared__ double s_part[256];
  const int64_t p0 = (int64_t)blockIdx.x * HP_GLISA_CHUNK;
  const int64_t p1 = p0 + HP_GLISA_CHUNK < npoint ? p0 + HP_GLISA_CHUNK : npoint;
  if (Mp <= 256) {
    const int ng = 256 / Mp;
    const int h = threadIdx.x / Mp, col = threadIdx.x - h * Mp;
    double s = 0.0;
    if (h < ng) {
#pragma unroll 8
      for (int64_t p = p0 + h; p < p1; p += ng) s += E[p] * G[p * Mp + col];
    }
    s_part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < M) {
      double v = 0.0;
      for (int g = 0; g < ng; ++g) v += s_part[g * Mp + threadIdx.x];
      grad_part[(int64_t)blockIdx.x * M + threadIdx.x] = v;
    }
  } else {
    for (int col = threadIdx.x; col < M; col += blockDim.x) {
      double s = 0.0;
#pragma unroll 8
      for (int64_t p = p0; p < p1; ++p) s += E[p] * G[p * Mp + col];
      grad_part[(int64_t)blockIdx.x * M + col] = s;
    }
  }
}

// Hessian tile partials, every tile pair of a point chunk (HP_GLISA_HCHUNK
// points) in one workgroup (Mp <= MPMAX; PW tile pairs per wavefront, at most
// 4 PW pairs).  Two workgroups share a compute unit, so that one's loads of G
// overlap the other's products.  The chunk's rows of G pass through LDS SC at
// a time (one contiguous run of SC Mp doubles, loaded into registers one
// round ahead).  Wavefront w accumulates
// the pairs w, w + 4, ... with the same rank-4 steps, in the same point
// order, as k_glisa_hess, reading each row of G once instead of once per
// tile pair.  With grad_part it also forms the gradient partials of the
// chunk from the staged rows (the even and the odd points of the chunk by
// two thread groups, then their sum), in place of k_glisa_grad's second pass
// over G.
template <int SC, int MPMAX, int PW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_glisa_hess_lds(
    int64_t npoint, int Mp, int npair, const int2 *__restrict__ pairs, const double *__restrict__ G,
    const double *__restrict__ D, double *__restrict__ hess_part, int M, const double *__restrict__ E,
    double *__restrict__ grad_part) {
  constexpr int RPT = SC * MPMAX / 256;
  __shared__ double s_g[SC * MPMAX];
  __shared__ double s_d[SC], s_e[SC];
  const int c = blockIdx.x, nchunk = gridDim.x;
  const int64_t p0 = (int64_t)c * HP_GLISA_HCHUNK;
  const int64_t p1 = p0 + HP_GLISA_HCHUNK < npoint ? p0 + HP_GLISA_HCHUNK : npoint;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, i = lane & 15, k = lane >> 4;
  int pI[PW], pJ[PW];
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    const int pi = wv + 4 * j;
    const int2 pr = pi < npair ? pairs[pi] : make_int2(0, 0);
    pI[j] = 16 * pr.x + i;
    pJ[j] = 16 * pr.y + i;
  }
  const int np = npair > wv ? (npair - wv + 3) / 4 : 0;  // pairs of this wavefront
  hp_gd4 acc[PW];
#pragma unroll
  for (int j = 0; j < PW; ++j) acc[j] = hp_gd4{0.0, 0.0, 0.0, 0.0};
  const int nround = (int)((p1 - p0 + SC - 1) / SC);
  double reg[RPT], dreg = 0.0, ereg = 0.0;
  const bool grad = grad_part != nullptr;
  // gradient: thread t < 2 Mp adds E g_c over the rows q = h, h + 2, ... of
  // every round (c = t mod Mp, h = t / Mp); the two halves are added in
  // order at the end
  const int gh = (int)threadIdx.x / Mp, gc = (int)threadIdx.x - gh * Mp;
  const bool gthread = grad && gh < 2 && gc < M;
  double gacc = 0.0;
  // round r's values: rows p0 + r SC ..., zero beyond the chunk
  auto fetch = [&](int r) {
    const int64_t pb = p0 + (int64_t)r * SC;
    const int64_t lim = (p1 - pb < SC ? p1 - pb : SC) * Mp;
    const double *src = G + pb * Mp;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      const int e = threadIdx.x + 256 * u;
      reg[u] = e < lim ? src[e] : 0.0;
    }
    const bool in = threadIdx.x < SC && pb + threadIdx.x < p1;
    dreg = in ? D[pb + threadIdx.x] : 0.0;
    ereg = (in && grad) ? E[pb + threadIdx.x] : 0.0;
  };
  fetch(0);
  for (int r = 0; r < nround; ++r) {
    __syncthreads();  // the previous round's rows are no longer read
#pragma unroll
    for (int u = 0; u < RPT; ++u) s_g[threadIdx.x + 256 * u] = reg[u];
    if (threadIdx.x < SC) {
      s_d[threadIdx.x] = dreg;
      s_e[threadIdx.x] = ereg;
    }
    __syncthreads();
    if (gthread)
      for (int q = gh; q < SC; q += 2) gacc += s_e[q] * s_g[q * Mp + gc];
    if (r + 1 < nround) fetch(r + 1);
    for (int q = 0; q < SC; q += 4) {
      const double dv = s_d[q + k];
      const double *row = s_g + (q + k) * Mp;
#pragma unroll
      for (int j = 0; j < PW; ++j) {
        if (j < np) {
          const double av = dv * row[pI[j]];
          const double bv = row[pJ[j]];
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[j], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    if (j < np) {
      double *out = hess_part + ((int64_t)(wv + 4 * j) * nchunk + c) * 256;
#pragma unroll
      for (int v = 0; v < 4; ++v) out[(4 * v + k) * 16 + i] = acc[j][v];
    }
  }
  if (grad) {
    __syncthreads();  // s_g is free: the halves meet there
    if (gthread) s_g[gh * Mp + gc] = gacc;
    __syncthreads();
    if ((int)threadIdx.x < M) grad_part[(int64_t)c * M + threadIdx.x] = s_g[threadIdx.x] + s_g[Mp + threadIdx.x];
  }
}

// Hessian tile partials: one wavefront per (tile pair I <= J, point chunk).
// Each step is a rank-4 update with 4 points: lane l supplies
// A[i][k] = D(p) G(p, 16 I + i) and B[k][j] = G(p, 16 J + j) with i = j = l % 16
// and p = p0 + k, k = l / 16.  Output layout (gfx950, FP64 16x16x4): lane l,
// register v holds C[4 v + l / 16][l % 16].
__global__ __launch_bounds__(64) void k_glisa_hess(int64_t npoint, int Mp, const int2 *__restrict__ pairs,
                                                   const double *__restrict__ G, const double *__restrict__ D,
                                                   double *__restrict__ hess_part) {
  const int2 pr = pairs[blockIdx.x];
  const int lane = threadIdx.x;
  const int i = lane & 15, k = lane >> 4;
  const int64_t p0 = (int64_t)blockIdx.y * HP_GLISA_HCHUNK;
  const int64_t p1 = p0 + HP_GLISA_HCHUNK < npoint ? p0 + HP_GLISA_HCHUNK : npoint;
  hp_gd4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int64_t q = p0; q < p1; q += 4) {
    const int64_t p = q + k;
    double av = 0.0, bv = 0.0;
    if (p < p1) {
      av = D[p] * G[p * Mp + 16 * pr.x + i];
      bv = G[p * Mp + 16 * pr.y + i];
    }
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  double *out = hess_part + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * 256;
#pragma unroll
  for (int v = 0; v < 4; ++v) out[(4 * v + k) * 16 + i] = acc[v];
}

// Fixed-order sums over the chunks: objective, gradient (negated) and the
// symmetric Hessian.  Workgroup 0 forms the objective: thread t adds the
// blocks of its contiguous range in block order, then thread 0 adds the 256
// ranges in order.  The other workgroups take one gradient or Hessian entry
// per thread (entry (blockIdx.x - 1) 256 + threadIdx.x), summed in chunk
// order with UB values loaded together before they are added, so the
// loads' latency is paid once per UB.
__global__ __launch_bounds__(256) void k_glisa_reduce(int nblock_obj, const double *__restrict__ obj_part, int nchunk,
                                                      int M, int do_grad, const double *__restrict__ grad_part, int npair,
                                                      const int2 *__restrict__ pairs, int hnchunk,
                                                      const double *__restrict__ hess_part, double *__restrict__ out) {
  // out: [0] objective, [1 .. 1+M) gradient, [1+M .. 1+M+M^2) Hessian, then
  // the smallest promolecule value
  constexpr int UB = 16;
  if (blockIdx.x == 0) {
    __shared__ double s_sum[256], s_min[256];
    const int seg = (nblock_obj + 255) / 256;
    const int b0 = threadIdx.x * seg, b1 = b0 + seg < nblock_obj ? b0 + seg : nblock_obj;
    double s = 0.0, mn = 1.0e308;
    for (int b = b0; b < b1; ++b) {
      s += obj_part[2 * b];
      mn = fmin(mn, obj_part[2 * b + 1]);
    }
    s_sum[threadIdx.x] = s;
    s_min[threadIdx.x] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0, tmn = 1.0e308;
      for (int u = 0; u < 256; ++u) {
        tot += s_sum[u];
        tmn = fmin(tmn, s_min[u]);
      }
      out[0] = tot;
      out[1 + M + (int64_t)M * M] = tmn;
    }
    return;
  }
  const int t = (blockIdx.x - 1) * blockDim.x + threadIdx.x;
  if (do_grad && t < M) {
    double s = 0.0;
    int c = 0;
    for (; c + UB <= nchunk; c += UB) {
      double v[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) v[u] = grad_part[(int64_t)(c + u) * M + t];
#pragma unroll
      for (int u = 0; u < UB; ++u) s += v[u];
    }
    for (; c < nchunk; ++c) s += grad_part[(int64_t)c * M + t];
    out[1 + t] = -s;
  }
  if (hess_part && t < npair * 256) {
    const int pi = t / 256, e = t % 256;
    const int2 pr = pairs[pi];
    const int ii = 16 * pr.x + e / 16, jj = 16 * pr.y + e % 16;
    if (ii < M && jj < M && ii <= jj) {
      const double *col = hess_part + (int64_t)pi * hnchunk * 256 + e;
      double s = 0.0;
      int c = 0;
      for (; c + UB <= hnchunk; c += UB) {
        double v[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) v[u] = col[(int64_t)(c + u) * 256];
#pragma unroll
        for (int u = 0; u < UB; ++u) s += v[u];
      }
      for (; c < hnchunk; ++c) s += col[(int64_t)c * 256];
      out[1 + M + (int64_t)ii * M + jj] = s;
      out[1 + M + (int64_t)jj * M + ii] = s;
    }
  }
}

// Per (point chunk, atom): the smallest value of the atom's pro-atom at cnew
// and sum_p w (rho0_a(cnew) - rho0_a(cold))^2 over the molecular grid.
template <int FK>
__global__ __launch_bounds__(256) void k_glisa_proatoms(DevMol m, DevFn fn, const double *__restrict__ norm,
                                                        const double *__restrict__ cnew,
                                                        const double *__restrict__ cold,
                                                        double *__restrict__ part) {
  __shared__ double lds[32];
  const int a = blockIdx.y;
  const int64_t p0 = (int64_t)blockIdx.x * HP_GLISA_CHUNK;
  const int64_t p1 = p0 + HP_GLISA_CHUNK < m.npoint ? p0 + HP_GLISA_CHUNK : m.npoint;
  double mn = 1.0e308, msd = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const double dx = m.px[p] - m.cx[a], dy = m.py[p] - m.cy[a], dz = m.pz[p] - m.cz[a];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    double y1 = 0.0, y2 = 0.0;
    for (int f = fn.fn_off[a]; f < fn.fn_off[a + 1]; ++f) {
      const double g = glisa_basis<FK>(fn, norm, f, r, r2);
      y1 += cnew[f] * g;
      y2 += cold[f] * g;
    }
    mn = fmin(mn, y1);
    const double d = y1 - y2;
    msd += m.wmol[p] * d * d;
  }
  const double lane_min = wave_min(mn);
  double dummy = 0.0;
  block_sum2(msd, dummy, lds);
  __shared__ double s_min[8];
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = lane_min;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = s_min[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmin(v, s_min[w]);
    part[((int64_t)a * gridDim.x + blockIdx.x) * 2] = v;
    part[((int64_t)a * gridDim.x + blockIdx.x) * 2 + 1] = msd;
  }
}

__global__ void k_glisa_proatoms_reduce(int natom, int nchunk, const double *__restrict__ part,
                                        double *__restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natom) return;
  constexpr int UB = 8;  // loaded together, added in chunk order
  const double *row = part + (int64_t)a * nchunk * 2;
  double mn = 1.0e308, msd = 0.0;
  int c = 0;
  for (; c + UB <= nchunk; c += UB) {
    double v[2 * UB];
#pragma unroll
    for (int u = 0; u < 2 * UB; ++u) v[u] = row[2 * c + u];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      mn = fmin(mn, v[2 * u]);
      msd += v[2 * u + 1];
    }
  }
  for (; c < nchunk; ++c) {
    mn = fmin(mn, row[2 * c]);
    msd += row[2 * c + 1];
  }
  out[2 * a] = mn;
  out[2 * a + 1] = msd;
}
