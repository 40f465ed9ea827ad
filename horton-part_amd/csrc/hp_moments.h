// Integrands of the AIM moments of any order (core/base.py:329-402).
//
// The quantities of order lmax, in the order of the output rows: the
// ncart = (L+1)(L+2)(L+3)/6 cartesian powers x^a y^b z^c (HORTON order: by l,
// then a from l down to 0, then b from l - a down to 0), the (L+1)^2 real
// regular solid harmonics (Racah normalisation, no Condon-Shortley phase;
// c_l0, c_l1, s_l1, c_l2, ...) and the radial powers r^0..r^L.
//
// Host and device code, so that the host can check the decoding and values.
#pragma once

#include <cmath>

#ifndef HP_HD
#ifdef __HIPCC__
#define HP_HD __host__ __device__
#else
#define HP_HD
#endif
#endif

struct MomQuantity {
  int kind;      // 0 cartesian, 1 solid harmonic, 2 radial, -1 none
  int i0, i1, i2;  // (a, b, c) | (l, m, sine) | (n, -, -)
  double norm;   // solid harmonics: sqrt(2 (l-m)!/(l+m)!) for m > 0, else 1
};

HP_HD inline MomQuantity decode_moment_quantity(int q, int lmax) {
  MomQuantity d{-1, 0, 0, 0, 1.0};
  const int ncart = (lmax + 1) * (lmax + 2) * (lmax + 3) / 6, npure = (lmax + 1) * (lmax + 1);
  if (q < 0) return d;
  if (q < ncart) {
    int l = 0;
    while ((l + 1) * (l + 2) * (l + 3) / 6 <= q) ++l;  // q's order: first index of order l+1 beyond q
    int k = q - l * (l + 1) * (l + 2) / 6;             // position within order l
    int a = l;
    while (k > l - a) {  // order l holds l - a + 1 powers with x^a
      k -= l - a + 1;
      --a;
    }
    d.kind = 0;
    d.i0 = a;
    d.i1 = l - a - k;
    d.i2 = k;
    return d;
  }
  q -= ncart;
  if (q < npure) {
    int l = 0;
    while ((l + 1) * (l + 1) <= q) ++l;
    const int k = q - l * l;
    const int m = (k + 1) / 2;
    d.kind = 1;
    d.i0 = l;
    d.i1 = m;
    d.i2 = (m > 0 && (k & 1) == 0) ? 1 : 0;
    if (m > 0) {
      double ratio = 1.0;  // (l-m)!/(l+m)!
      for (int t = l - m + 1; t <= l + m; ++t) ratio /= (double)t;
      d.norm = std::sqrt(2.0 * ratio);
    }
    return d;
  }
  q -= npure;
  if (q <= lmax) {
    d.kind = 2;
    d.i0 = q;
  }
  return d;
}

// Value of one quantity at the offset (x, y, z).  Solid harmonics from
// (x + iy)^m and the upward recurrence in l at fixed m of
// P(l, m) = r^(l-m) d^m P_l(cos t)/d(cos t)^m, a polynomial in z and r^2:
// P(m, m) = (2m-1)!!, P(m+1, m) = (2m+1) z P(m, m),
// P(l, m) = ((2l-1) z P(l-1, m) - (l+m-1) r^2 P(l-2, m)) / (l-m).
HP_HD inline double moment_quantity(const MomQuantity &d, double x, double y, double z) {
  if (d.kind == 0) {
    double v = 1.0;
    for (int k = 0; k < d.i0; ++k) v *= x;
    for (int k = 0; k < d.i1; ++k) v *= y;
    for (int k = 0; k < d.i2; ++k) v *= z;
    return v;
  }
  const double r2 = x * x + y * y + z * z;
  if (d.kind == 1) {
    const int l = d.i0, m = d.i1;
    double c = 1.0, s = 0.0;
    for (int k = 0; k < m; ++k) {
      const double cn = c * x - s * y;
      s = c * y + s * x;
      c = cn;
    }
    double pmm = 1.0;
    for (int k = 3; k < 2 * m; k += 2) pmm *= (double)k;
    double pl = pmm;
    if (l > m) {
      double prev = pmm;
      pl = (double)(2 * m + 1) * z * pmm;
      for (int ll = m + 2; ll <= l; ++ll) {
        const double nx = ((double)(2 * ll - 1) * z * pl - (double)(ll + m - 1) * r2 * prev) / (double)(ll - m);
        prev = pl;
        pl = nx;
      }
    }
    return d.norm * pl * (m == 0 ? 1.0 : (d.i2 ? s : c));
  }
  if (d.kind == 2) {
    const double r = std::sqrt(r2);
    double v = 1.0;
    for (int k = 0; k < d.i0; ++k) v *= r;
    return v;
  }
  return 0.0;
}

