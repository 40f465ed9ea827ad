// Piecewise cubic radial tables of the tabulated pro-atoms.
//
// A table over ascending knots x_0 < ... < x_{n-1} holds, for each interval
// i < n - 1, the cubic c0 d^3 + c1 d^2 + c2 d + c3 with d = r - x_i, as four
// doubles at coef[4 i .. 4 i + 3] (row n - 1 is zero).  This is scipy's PPoly
// layout (c[0] the cubic term), which the reference's splines use:
// CubicSpline for densities without derivatives and CubicHermiteSpline for
// those with (core/stockholder.py:227-264, core/proatomdb.py:378-388), both
// with extrapolation beyond the end knots (the end pieces continue).
//
// Host and device code: the ISA update builds its not-a-knot tables on the
// device every iteration.
#pragma once

#ifndef HP_HD
#ifdef __HIPCC__
#define HP_HD __host__ __device__
#else
#define HP_HD
#endif
#endif

// Hermite piece of interval [x0, x0 + h] with end values y0, y1 and end
// slopes s0, s1 (scipy CubicHermiteSpline): t = (s0 + s1 - 2 m) / h with
// m = (y1 - y0) / h, c0 = t / h, c1 = (m - s0) / h - t, c2 = s0, c3 = y0.
HP_HD inline void hermite_piece(double h, double y0, double y1, double s0, double s1, double *c) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double m = (y1 - y0) / h;
  const double t = (s0 + s1 - 2.0 * m) / h;
  c[0] = t / h;
  c[1] = (m - s0) / h - t;
  c[2] = s0;
  c[3] = y0;
}

// Not-a-knot cubic spline through (x_i, y_i), n >= 4 (scipy CubicSpline's
// default boundary): the knot slopes s_i solve the tridiagonal system
//   row 0:      dx1 s0 + (x2 - x0) s1 = ((dx0 + 2 (x2 - x0)) dx1 m0 + dx0^2 m1) / (x2 - x0)
//   row i:      dx_i s_{i-1} + 2 (dx_{i-1} + dx_i) s_i + dx_{i-1} s_{i+1}
//                   = 3 (dx_i m_{i-1} + dx_{i-1} m_i)
//   row n-1:    (x_{n-1} - x_{n-3}) s_{n-2} + dx_{n-3} s_{n-1}
//                   = (dx_{n-2}^2 m_{n-3} + (2 (x_{n-1} - x_{n-3}) + dx_{n-2}) dx_{n-3} m_{n-2}) / (x_{n-1} - x_{n-3})
// (dx_i = x_{i+1} - x_i, m_i = (y_{i+1} - y_i) / dx_i; the first and last rows
// make the third derivative continuous at x_1 and x_{n-2}), eliminated in
// order (Thomas); the pieces are then the Hermite cubics of the slopes.  cp
// and rp are scratch arrays of n values; x, y, cp and rp may live in LDS.
HP_HD inline void notaknot_spline(int n, const double *x, const double *y, double *cp, double *rp, double *coef) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  {
    const double dx0 = x[1] - x[0], dx1 = x[2] - x[1], d = x[2] - x[0];
    const double m0 = (y[1] - y[0]) / dx0, m1 = (y[2] - y[1]) / dx1;
    const double r = ((dx0 + 2.0 * d) * dx1 * m0 + dx0 * dx0 * m1) / d;
    cp[0] = d / dx1;
    rp[0] = r / dx1;
  }
  double dxm = x[1] - x[0], mm = (y[1] - y[0]) / dxm;
  for (int i = 1; i < n - 1; ++i) {
    const double dxi = x[i + 1] - x[i];
    const double mi = (y[i + 1] - y[i]) / dxi;
    const double r = 3.0 * (dxi * mm + dxm * mi);
    const double piv = 2.0 * (dxm + dxi) - dxi * cp[i - 1];
    cp[i] = dxm / piv;
    rp[i] = (r - dxi * rp[i - 1]) / piv;
    dxm = dxi;
    mm = mi;
  }
  {
    const double dxl = x[n - 1] - x[n - 2], dxp = x[n - 2] - x[n - 3], d = x[n - 1] - x[n - 3];
    const double ml = (y[n - 1] - y[n - 2]) / dxl, mp = (y[n - 2] - y[n - 3]) / dxp;
    const double r = (dxl * dxl * mp + (2.0 * d + dxl) * dxp * ml) / d;
    const double piv = dxp - d * cp[n - 2];
    rp[n - 1] = (r - d * rp[n - 2]) / piv;
  }
  double s1 = rp[n - 1];
  coef[4 * (n - 1) + 0] = 0.0;
  coef[4 * (n - 1) + 1] = 0.0;
  coef[4 * (n - 1) + 2] = 0.0;
  coef[4 * (n - 1) + 3] = 0.0;
  for (int i = n - 2; i >= 0; --i) {
    const double s0 = rp[i] - cp[i] * s1;
    hermite_piece(x[i + 1] - x[i], y[i], y[i + 1], s0, s1, coef + 4 * i);
    s1 = s0;
  }
}

// Interval of r in the knots (scipy PPoly's search): the largest i <= n - 2
// with x_i <= r, or 0 below x_0 (r beyond x_{n-1} lands in n - 2).
HP_HD inline int table_interval(const double *x, int n, double r) {
  int lo = 0, hi = n - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (x[mid] <= r)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Value of a piece at d = r - x_i: c3 + c2 d + c1 d^2 + c0 d^3 with the powers of d
// built by repeated products and every product and sum rounded on its own
// (the operations of scipy's PPoly evaluation, without fused multiply-adds).
// Across a wide last interval a density's cubic cancels many digits, so the
// order of operations shows in its far tail.
HP_HD inline double table_piece(const double *c, double d) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  double z = d;
  double v = c[3] + c[2] * z;
  z *= d;
  v += c[1] * z;
  z *= d;
  v += c[0] * z;
  return v;
}

HP_HD inline double table_eval(const double *x, const double *coef, int n, double r) {
  const int i = table_interval(x, n, r);
  return table_piece(coef + 4 * i, r - x[i]);
}
