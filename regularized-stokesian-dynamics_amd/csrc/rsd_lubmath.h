// rsd_lubmath.h -- two-sphere lubrication scalars and their projected blocks.
//
// lubcalc restates hydro_lub.f90:310-398 (near-contact closed forms for
// s < 2.1) and :400-525 (linear interpolation in the r2babc / r2bgh / r2bm
// tables with the bracket t(i) <= s < t(i+1), zero outside, mobility-m
// entries halved; the bracket is found by bisection).  lub_blocks evaluates the Y2-projected blocks of
// hydro_lub.f90:111-176 in closed form:
//   a = X e e + Y (delta - e e),  b = Y eps.e,
//   g_Y(I,k) = (Xg - 2Yg)(e.Y_I e) e_k + 2 Yg (Y_I e)_k,   h_Y(I,k) = 2 Yh (e x Y_I e)_k,
//   m(I,J) = 3/2 Xm (eY_Ie)(eY_Je) + 2 Ym (Y_Ie.Y_Je - eY_Ie eY_Je)
//            + Zm (d_IJ - 2 Y_Ie.Y_Je + 1/2 eY_Ie eY_Je).
#pragma once
#include "rsd_pairmath.h"

namespace rsdm {

constexpr int kLubAbcRows = 39, kLubGhRows = 47, kLubMRows = 47;

struct LubTables {
  double abc[kLubAbcRows * 11];  // rows as in r2babc.dat: r, X11A X12A Y11A Y12A Y11B Y12B X11C X12C Y11C Y12C
  double gh[kLubGhRows * 7];     // r, X11G X12G Y11G Y12G Y11H Y12H
  double m[kLubMRows * 4];       // r, X11M+X12M, Y11M+Y12M, Z11M+Z12M
};

// Row i with t(i) <= r < t(i+1) in a table of strictly increasing first
// columns (row stride `cols`), or -1: the bracket of hydro_lub.f90:488-523,
// found by bisection instead of a scan (same row, fewer divergent steps).
RSD_HD int lub_bracket(const double* tab, int cols, int rows, double r) {
  if (!(tab[0] <= r && r < tab[(rows - 1) * cols])) return -1;
  int lo = 0, hi = rows - 1;  // tab[lo] <= r < tab[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tab[mid * cols] <= r) lo = mid;
    else hi = mid;
  }
  return lo;
}

// v[19] = x11a,x12a,y11a,y12a,y11b,y12b,x11c,x12c,y11c,y12c,x11g,x12g,y11g,y12g,y11h,y12h,xm,ym,zm
RSD_HD void lubcalc(double r, const LubTables& t, double* v) {
  if (r < 2.1) {
    const double xi = r - 2.0, invxi = 1.0 / xi, li = log(invxi), xl = xi * li;
    v[0] = -1.23041 + 0.25 * invxi + 1.8918 * xi + 9.0 * li / 40.0 + 3.0 * xl / 112.0;
    v[1] = -v[0] + 0.00312 - 0.0011 * xi;
    v[2] = -0.39394 + 0.95665 * xi + li / 6.0;
    v[3] = -v[2] + 0.004636 - 0.007049 * xi;
    v[4] = 0.408286 - 0.84055 * xi - li / 6.0 - xl / 12.0;
    v[5] = -0.405978 + 0.833042 * xi + li / 6.0 + xl / 12.0;
    v[6] = 0.0479 + 0.12494 * xi - xl / 6.0;
    v[7] = -v[6] + 0.016869 + 0.049536 * xi;
    v[8] = -0.605434 + 0.939139 * xi + 4.0 * li / 15.0 + 94.0 * xl / 375.0;
    v[9] = li / 15.0 + 62.0 * xl / 375.0 - 0.219032 + 0.332843 * xi;
    v[10] = -1.16897 + 1.47882 * xi + invxi / 4.0 + 9.0 * li / 40.0 + 39.0 * xl / 280.0;
    v[11] = -v[10] + 0.01 - 0.00167 * xi;
    v[12] = -0.2041 + 0.44226 * xi + li / 12.0 + xl / 24.0;
    v[13] = -v[12] + 0.012265 - 0.02757 * xi;
    v[14] = -0.143777 + 0.264207 * xi + li / 30.0 + 137.0 * xl / 1500.0;
    v[15] = -0.298166 + 0.534123 * xi + 2.0 * li / 15.0 + 113.0 * xl / 1500.0;
    v[16] = invxi / 6.0 + 3.0 * li / 20.0 + 47.0 * xl / 280.0 - 0.740815 + 0.706802 * xi;
    v[17] = -0.193518 + 0.553883 * xi + li / 12.0 - 29.0 * xl / 500.0;
    v[18] = 0.00645755 - 0.021142 * xi;
    return;
  }
  for (int q = 0; q < 19; ++q) v[q] = 0.0;
  int i = lub_bracket(t.abc, 11, kLubAbcRows, r);
  if (i >= 0) {
    const double t0 = t.abc[i * 11], t1 = t.abc[(i + 1) * 11];
    for (int c = 0; c < 10; ++c) {
      const double y0 = t.abc[i * 11 + 1 + c], y1 = t.abc[(i + 1) * 11 + 1 + c];
      v[c] = y0 + (r - t0) * (y1 - y0) / (t1 - t0);
    }
  }
  i = lub_bracket(t.gh, 7, kLubGhRows, r);
  if (i >= 0) {
    const double t0 = t.gh[i * 7], t1 = t.gh[(i + 1) * 7];
    for (int c = 0; c < 6; ++c) {
      const double y0 = t.gh[i * 7 + 1 + c], y1 = t.gh[(i + 1) * 7 + 1 + c];
      v[10 + c] = y0 + (r - t0) * (y1 - y0) / (t1 - t0);
    }
  }
  i = lub_bracket(t.m, 4, kLubMRows, r);
  if (i >= 0) {
    const double t0 = t.m[i * 4], t1 = t.m[(i + 1) * 4];
    for (int c = 0; c < 3; ++c) {
      const double y0 = t.m[i * 4 + 1 + c], y1 = t.m[(i + 1) * 4 + 1 + c];
      v[16 + c] = (y0 + (r - t0) * (y1 - y0) / (t1 - t0)) / 2.0;
    }
  }
}

// One pair's projected blocks for a given orientation e.  Symmetric 3x3 in
// the order 00,11,22,01,02,12; b as the vector b_k with B(i,j) = eps_ijk b_k;
// g, h as (I,c) at [I + 5c]; m packed upper 5x5.
struct LubBlock {
  double a[6], c[6], b[3], g[15], h[15];
};

RSD_HD void lub_block(const double* e, double X_a, double Y_a, double Y_b, double X_c, double Y_c, double X_g,
                      double Y_g, double Y_h, const YProj& p, LubBlock& o) {
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    const double ee = e[si[q]] * e[sj[q]], kd = (q < 3) ? 1.0 : 0.0;
    o.a[q] = X_a * ee + Y_a * (kd - ee);
    o.c[q] = X_c * ee + Y_c * (kd - ee);
  }
  for (int k = 0; k < 3; ++k) o.b[k] = Y_b * e[k];
  for (int I = 0; I < 5; ++I) {
    const double* v = p.Yx[I];
    const double cx[3] = {e[1] * v[2] - e[2] * v[1], e[2] * v[0] - e[0] * v[2], e[0] * v[1] - e[1] * v[0]};
    for (int k = 0; k < 3; ++k) {
      o.g[I + 5 * k] = (X_g - 2.0 * Y_g) * p.xYx[I] * e[k] + 2.0 * Y_g * v[k];
      o.h[I + 5 * k] = 2.0 * Y_h * cx[k];
    }
  }
}

RSD_HD void lub_mmx(double xm, double ym, double zm, const YProj& p, double* mm) {
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k) {
      const double yy = p.Yx[I][0] * p.Yx[J][0] + p.Yx[I][1] * p.Yx[J][1] + p.Yx[I][2] * p.Yx[J][2];
      const double ee = p.xYx[I] * p.xYx[J];
      mm[k] = 1.5 * xm * ee + 2.0 * ym * (yy - ee) + zm * ((I == J ? 1.0 : 0.0) - 2.0 * yy + 0.5 * ee);
    }
}

}  // namespace rsdm
