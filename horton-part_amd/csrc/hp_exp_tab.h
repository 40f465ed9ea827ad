// FP64 exponential of non-positive arguments by Tang's table-driven method
// (P. T. P. Tang, ACM Trans. Math. Softw. 15 (1989) 144-157), shared by the
// device kernels and a host-side accuracy test.
//
// x = (64 m + j) ln2/64 + r with |r| <= ln2/128;
// exp(x) = 2^m * 2^(j/64) * (1 + p(r)),  p(r) = r + r^2/2 + ... + r^5/120.
// The truncation of p is below 4e-17 relative; ln2/64 is split into two
// doubles for the reduction; the 64 table values 2^(j/64) are correctly
// rounded.  Arguments below -1000 give 0 (exp underflows at -745.13).
#pragma once

#ifndef HP_HD
#define HP_HD
#endif

// 2^(j/64), j = 0..63, correctly rounded
#define HP_EXP2_64_VALUES \
  { \
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0, \
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, \
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0, \
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0, \
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, \
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0, \
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0, \
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, \
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0, \
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0, \
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, \
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0, \
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0, \
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, \
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0, \
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0  \
  }

HP_HD inline double hp_exp_tab_core(double x, const double *t64) {
  x = fmax(x, -1000.0);  // keeps 64 m + j in int range
  const double k = rint(x * 0x1.71547652b82fep+6);  // 64 / ln2
  double r = fma(-k, 0x1.62e42fefa39efp-7, x);     // ln2/64, high part
  r = fma(-k, 0x1.abc9e3b39803fp-62, r);           // ln2/64, low part
  const int ki = (int)k;
  const double t = t64[ki & 63];
  double p = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p *= r;
  return ldexp(fma(t, p, t), ki >> 6);  // ki >> 6 floors: m = floor(k / 64)
}

// 2^(y/64) for y <= 0: the same table method with the argument already in
// units of ln2/64, as the weights pass forms it (y = -(alpha 64/ln2) r, the
// exponent scaled once per function).  With k = rint(y) the reduced argument
// f = y - k is exact, |f| <= 1/2, and e^(f ln2/64) - 1 is the degree-5 Taylor
// polynomial in f with coefficients (ln2/64)^n / n! (truncation below 4e-17
// relative, as above).  Arguments below -1000 64/ln2 give 0.
HP_HD inline double hp_exp2_64th_core(double y, const double *t64) {
  y = fmax(y, -92332.48261689366);  // -1000 * 64/ln2: keeps k in int range
  const double k = rint(y);
  const double f = y - k;
  const int ki = (int)k;
  const double t = t64[ki & 63];
  double p = fma(f, 0x1.5d87fe78a6731p-40, 0x1.3b2ab6fba4e77p-31);
  p = fma(p, f, 0x1.c6b08d704a0c0p-23);
  p = fma(p, f, 0x1.ebfbdff82c58fp-15);
  p = fma(p, f, 0x1.62e42fefa39efp-7);
  p *= f;
  return ldexp(fma(t, p, t), ki >> 6);  // ki >> 6 floors: m = floor(k / 64)
}

// 64 / ln2, rounded to the nearest double: exponents of the weights pass are
// stored multiplied by it
#define HP_64_OVER_LN2 0x1.71547652b82fep+6
