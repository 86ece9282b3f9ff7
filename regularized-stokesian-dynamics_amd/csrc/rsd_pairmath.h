// rsd_pairmath.h -- per-image and per-wavevector RPY grand-mobility terms in
// isotropic closed form.
//
// The reference (Ewald_summation.f90:196-423, Regularization.f90:18-162,
// stokesian_Green.f90:11-108) builds every pair block from full Cartesian
// derivative tensors of a radial function f(r) (up to sixth order, 3^4
// arrays) and then projects them onto the Y2 basis (tensors.f90:43-67,
// 936-1023).  All those tensors are isotropic: a handful of scalar radial
// functions times {delta, x x, x x x x, eps.x}.  This header evaluates the
// projected 11x11 block directly from those scalars:
//
//   g_n  = ((1/r) d/dr)^n f            h = lap f     q = lap^2 f
//   D^n h, D^n q likewise.
//
// The three radial families are f = r erfc(lambda r) (Ewald real space,
// helpers at Ewald_summation.f90:658-1053), f' = erf(kappa r) (regularised
// overlap kernel, Regularization.f90:400-778) and f = r (free-space Oseen,
// stokesian_Green.f90:215-434).  The free-space helper Init_DDD2_Jg carries a
// sign flip in its delta_ij x_k x_l term relative to the exact derivative;
// its effect on the projected blocks is added explicitly (oseen_quirk).
//
// Output layout (60 doubles, PairOut):
//   tt[6], rr[6]  symmetric 3x3 in the order 00,11,22,01,02,12
//   tr[3]         TR(i,j) = eps_ijm tr[m]
//   gq[15], hq[15] 3x5, element (i,I) at [i + 3*I]  (GH_TR layout)
//   qq[15]        symmetric 5x5, packed upper: (I,J) I<=J row by row
#pragma once
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define RSD_HD __host__ __device__ __forceinline__
#else
#define RSD_HD static inline
#endif

namespace rsdm {

struct PairOut {
  double tt[6], rr[6], tr[3], gq[15], hq[15], qq[15];
};

RSD_HD void pairout_zero(PairOut& o) {
  for (int i = 0; i < 6; ++i) o.tt[i] = o.rr[i] = 0.0;
  for (int i = 0; i < 3; ++i) o.tr[i] = 0.0;
  for (int i = 0; i < 15; ++i) o.gq[i] = o.hq[i] = o.qq[i] = 0.0;
}

RSD_HD int qq_index(int I, int J) {  // I <= J
  return I * 5 - I * (I - 1) / 2 + (J - I);
}

// Scalars of one radial family at one separation.
struct Radial {
  double hg1, g2, g3, g4;         // h - g1, D^2 f, D^3 f, D^4 f
  double Dh, D2h, D3h, D4h;       // D^n h
  double qDh, Dq, D2q;            // q - Dh, D q, D^2 q
};

// f = r erfc(lam r): e = erfc(lam r), e1..e6 its r-derivatives (Ewald_summation.f90:658-743)
RSD_HD void radial_ewald(double r, double lam, Radial& s) {
  const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir, ir6 = ir3 * ir3,
               ir7 = ir6 * ir, ir8 = ir4 * ir4, ir9 = ir8 * ir;
  const double l2 = lam * lam, rl2 = r * r * l2, rl4 = rl2 * rl2;
  const double P = 1.1283791670955125738961589031215452 * lam * exp(-rl2);  // 2 lam/sqrt(pi) exp(-l^2 r^2)
  const double E = erfc(lam * r);
  const double E1 = -P;
  const double E2 = 2.0 * l2 * r * P;
  const double E3 = -2.0 * l2 * P * (2.0 * rl2 - 1.0);
  const double E4 = 4.0 * l2 * l2 * r * P * (2.0 * rl2 - 3.0);
  const double E5 = -4.0 * l2 * l2 * P * (4.0 * rl4 - 12.0 * rl2 + 3.0);
  const double E6 = 8.0 * l2 * l2 * l2 * r * P * (4.0 * rl4 - 20.0 * rl2 + 15.0);
  s.hg1 = E * ir + 3.0 * E1 + r * E2;
  s.g2 = -E * ir3 + E1 * ir2 + E2 * ir;
  s.g3 = 3.0 * E * ir5 - 3.0 * E1 * ir4 + E3 * ir2;
  s.g4 = -15.0 * E * ir7 + 15.0 * E1 * ir6 - 3.0 * E2 * ir5 - 2.0 * E3 * ir4 + E4 * ir3;
  s.Dh = -2.0 * E * ir3 + 2.0 * E1 * ir2 + 5.0 * E2 * ir + E3;
  s.D2h = 6.0 * E * ir5 - 6.0 * E1 * ir4 - 3.0 * E2 * ir3 + 5.0 * E3 * ir2 + E4 * ir;
  s.D3h = -30.0 * E * ir7 + 30.0 * E1 * ir6 + 3.0 * E2 * ir5 - 13.0 * E3 * ir4 + 4.0 * E4 * ir3 + E5 * ir2;
  s.D4h = 210.0 * E * ir9 - 210.0 * E1 * ir8 + 15.0 * E2 * ir7 + 55.0 * E3 * ir6 - 25.0 * E4 * ir5 + 2.0 * E5 * ir4 +
          E6 * ir3;
  const double q = 12.0 * E2 * ir + 8.0 * E3 + r * E4;
  s.qDh = q - s.Dh;
  s.Dq = -12.0 * E2 * ir3 + 12.0 * E3 * ir2 + 9.0 * E4 * ir + E5;
  s.D2q = 36.0 * E2 * ir5 - 36.0 * E3 * ir4 + 3.0 * E4 * ir3 + 9.0 * E5 * ir2 + E6 * ir;
}

// f' = erf(kap r) (Regularization.f90:400-472)
RSD_HD void radial_regular(double r, double kap, Radial& s) {
  const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir, ir6 = ir3 * ir3,
               ir7 = ir6 * ir, ir8 = ir4 * ir4, ir9 = ir8 * ir;
  const double k2 = kap * kap, rl2 = r * r * k2, rl4 = rl2 * rl2;
  const double E = 1.1283791670955125738961589031215452 * kap * exp(-rl2);
  const double P = erf(kap * r);
  const double P1 = E;
  const double P2 = -2.0 * k2 * r * E;
  const double P3 = 2.0 * k2 * E * (2.0 * rl2 - 1.0);
  const double P4 = -4.0 * k2 * k2 * r * E * (2.0 * rl2 - 3.0);
  const double P5 = 4.0 * k2 * k2 * E * (4.0 * rl4 - 12.0 * rl2 + 3.0);
  s.hg1 = P * ir + P1;
  s.g2 = -P * ir3 + P1 * ir2;
  s.g3 = 3.0 * P * ir5 - 3.0 * P1 * ir4 + P2 * ir3;
  s.g4 = -15.0 * P * ir7 + 15.0 * P1 * ir6 - 6.0 * P2 * ir5 + P3 * ir4;
  s.Dh = -2.0 * P * ir3 + 2.0 * P1 * ir2 + P2 * ir;
  s.D2h = 6.0 * P * ir5 - 6.0 * P1 * ir4 + P2 * ir3 + P3 * ir2;
  s.D3h = -30.0 * P * ir7 + 30.0 * P1 * ir6 - 9.0 * P2 * ir5 - P3 * ir4 + P4 * ir3;
  s.D4h = 210.0 * P * ir9 - 210.0 * P1 * ir8 + 75.0 * P2 * ir7 - 5.0 * P3 * ir6 - 4.0 * P4 * ir5 + P5 * ir4;
  const double q = 4.0 * P2 * ir + P3;
  s.qDh = q - s.Dh;
  s.Dq = -4.0 * P2 * ir3 + 4.0 * P3 * ir2 + P4 * ir;
  s.D2q = 12.0 * P2 * ir5 - 12.0 * P3 * ir4 + 3.0 * P4 * ir3 + P5 * ir2;
}

// f = r (Oseen; stokesian_Green.f90:215-296 without the DDD2 quirk)
RSD_HD void radial_oseen(double r, Radial& s) {
  const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir5 = ir3 * ir2, ir7 = ir5 * ir2, ir9 = ir7 * ir2;
  s.hg1 = ir;
  s.g2 = -ir3;
  s.g3 = 3.0 * ir5;
  s.g4 = -15.0 * ir7;
  s.Dh = -2.0 * ir3;
  s.D2h = 6.0 * ir5;
  s.D3h = -30.0 * ir7;
  s.D4h = 210.0 * ir9;
  s.qDh = 2.0 * ir3;
  s.Dq = 0.0;
  s.D2q = 0.0;
}

// Y2 basis contractions of a vector x (tensors.f90:43-67)
struct YProj {
  double Yx[5][3];  // Y_I x
  double xYx[5];    // x . Y_I x
};
RSD_HD void yproj(double x, double y, double z, YProj& p) {
  const double s2 = 0.70710678118654752440084436210484903, s6 = 0.40824829046386301636621401245098189;
  p.Yx[0][0] = s2 * y;
  p.Yx[0][1] = s2 * x;
  p.Yx[0][2] = 0.0;
  p.Yx[1][0] = 0.0;
  p.Yx[1][1] = -s2 * z;
  p.Yx[1][2] = -s2 * y;
  p.Yx[2][0] = -s6 * x;
  p.Yx[2][1] = -s6 * y;
  p.Yx[2][2] = 2.0 * s6 * z;
  p.Yx[3][0] = -s2 * z;
  p.Yx[3][1] = 0.0;
  p.Yx[3][2] = -s2 * x;
  p.Yx[4][0] = s2 * x;
  p.Yx[4][1] = -s2 * y;
  p.Yx[4][2] = 0.0;
  for (int I = 0; I < 5; ++I) p.xYx[I] = x * p.Yx[I][0] + y * p.Yx[I][1] + z * p.Yx[I][2];
}

// Accumulate w * (projected real-space block of one radial family at x).
//   ai2, aj2: squared radii; phix: (1 - phi/5) on the a^2 Laplacian terms of TT
//   (Ewald only; 1 otherwise); tr_gh: whether TR/GQ/HQ are wanted (not for self).
RSD_HD void radial_block(double x, double y, double z, const Radial& s, double ai2, double aj2, double phix,
                         bool tr_gh, double w, PairOut& o) {
  const double ct = (ai2 + aj2) / 6.0 * phix;
  // TT = (1/8)[J + ct lap J],  J = delta (h-g1) - xx g2,  lap J = delta (q-Dh) - xx D2h
  {
    const double dcoef = w * 0.125 * (s.hg1 + ct * s.qDh);
    const double xcoef = -w * 0.125 * (s.g2 + ct * s.D2h);
    o.tt[0] += dcoef + xcoef * x * x;
    o.tt[1] += dcoef + xcoef * y * y;
    o.tt[2] += dcoef + xcoef * z * z;
    o.tt[3] += xcoef * x * y;
    o.tt[4] += xcoef * x * z;
    o.tt[5] += xcoef * y * z;
  }
  // RR = -(1/32) lap J
  {
    const double dcoef = -w * (1.0 / 32.0) * s.qDh;
    const double xcoef = w * (1.0 / 32.0) * s.D2h;
    o.rr[0] += dcoef + xcoef * x * x;
    o.rr[1] += dcoef + xcoef * y * y;
    o.rr[2] += dcoef + xcoef * z * z;
    o.rr[3] += xcoef * x * y;
    o.rr[4] += xcoef * x * z;
    o.rr[5] += xcoef * y * z;
  }
  YProj p;
  yproj(x, y, z, p);
  if (tr_gh) {
    const double tc = -w * (1.0 / 16.0) * (s.Dh + ai2 / 6.0 * s.Dq);
    o.tr[0] += tc * x;
    o.tr[1] += tc * y;
    o.tr[2] += tc * z;
    const double cg = ai2 / 6.0 + 0.1 * aj2;
    const double gA = -w * 0.125 * (s.Dh - 2.0 * s.g2 + cg * (s.Dq - 2.0 * s.D2h));
    const double gB = w * 0.125 * (s.g3 + cg * s.D3h);
    const double hC = -w * (1.0 / 16.0) * (s.D2h + 0.1 * aj2 * s.D2q);
    const double xv[3] = {x, y, z};
    for (int I = 0; I < 5; ++I) {
      const double* v = p.Yx[I];
      const double cx[3] = {y * v[2] - z * v[1], z * v[0] - x * v[2], x * v[1] - y * v[0]};  // x cross Yx
      for (int i = 0; i < 3; ++i) {
        o.gq[i + 3 * I] += gA * v[i] + gB * xv[i] * p.xYx[I];
        o.hq[i + 3 * I] += hC * cx[i];
      }
    }
  }
  // QQ = -(1/8) PA - ((ai2+aj2)/80) PB with
  //   PA = d_IJ (Dh-2g2) + (Yx_I.Yx_J)(D2h-4g3) - xYx_I xYx_J g4
  //   PB = d_IJ (Dq-2D2h) + (Yx_I.Yx_J)(D2q-4D3h) - xYx_I xYx_J D4h
  {
    const double c80 = (ai2 + aj2) / 80.0;
    const double cd = -w * (0.125 * (s.Dh - 2.0 * s.g2) + c80 * (s.Dq - 2.0 * s.D2h));
    const double cy = -w * (0.125 * (s.D2h - 4.0 * s.g3) + c80 * (s.D2q - 4.0 * s.D3h));
    const double cx = w * (0.125 * s.g4 + c80 * s.D4h);
    int k = 0;
    for (int I = 0; I < 5; ++I)
      for (int J = I; J < 5; ++J, ++k) {
        const double yy = p.Yx[I][0] * p.Yx[J][0] + p.Yx[I][1] * p.Yx[J][1] + p.Yx[I][2] * p.Yx[J][2];
        o.qq[k] += (I == J ? cd : 0.0) + cy * yy + cx * p.xYx[I] * p.xYx[J];
      }
  }
}

// The effect of the sign flip in Init_DDD2_Jg (stokesian_Green.f90:296-311) on
// the free-space blocks: HQ += 0.375 aj^2 / r^7 (x cross Yx), QQ += 0.75 (ai^2+aj^2)/r^7 (Yx_I . Yx_J).
RSD_HD void oseen_quirk(double x, double y, double z, double ai2, double aj2, double w, PairOut& o) {
  const double r2 = x * x + y * y + z * z, r = sqrt(r2);
  const double ir7 = 1.0 / (r2 * r2 * r2 * r);
  YProj p;
  yproj(x, y, z, p);
  const double hc = w * 0.375 * aj2 * ir7, qc = w * 0.75 * (ai2 + aj2) * ir7;
  for (int I = 0; I < 5; ++I) {
    const double* v = p.Yx[I];
    const double cx[3] = {y * v[2] - z * v[1], z * v[0] - x * v[2], x * v[1] - y * v[0]};
    for (int i = 0; i < 3; ++i) o.hq[i + 3 * I] += hc * cx[i];
  }
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k)
      o.qq[k] += qc * (p.Yx[I][0] * p.Yx[J][0] + p.Yx[I][1] * p.Yx[J][1] + p.Yx[I][2] * p.Yx[J][2]);
}

// Overlap correction (GRPY_YZD.f90:813-836): regularised RPY minus free-space RPY,
// kappa = sqrt(pi)/(3 |ai+aj-r|) (Regularization.f90:69).
RSD_HD void overlap_regular_correction(double x, double y, double z, double ai, double aj, PairOut& o) {
  const double r = sqrt(x * x + y * y + z * z);
  const double kap = 1.7724538509055160272981674833411452 / 3.0 / fabs((ai + aj) - r);
  Radial s;
  radial_regular(r, kap, s);
  radial_block(x, y, z, s, ai * ai, aj * aj, 1.0, true, 1.0, o);
  radial_oseen(r, s);
  radial_block(x, y, z, s, ai * ai, aj * aj, 1.0, true, -1.0, o);
  oseen_quirk(x, y, z, ai * ai, aj * aj, -1.0, o);
}

// Rotne-Prager-Yamakawa blocks of an overlapping pair (YAMAKAWA_IJ,
// GRPY_YZD.f90:326-786), accumulated with weight w.  The reference builds
// Cartesian tensors from t = r/|r| (delta t, t t t, eps.t t and the fourth-order
// D0, D1, D2 of tensors.f90:347-440) and projects them on the Y2 basis; here the
// projections are taken in closed form, with v_I = Y_I t and p_I = t . Y_I t:
//   GQ(i,I) = cU v_I,i + cR t_i p_I             (t3UR, t3RRR)
//   HQ(i,I) = cH (v_I x t)_i                     (eps_ijm t_m t_k)
//   QQ(I,J) = c2 d_IJ + 2 (c1 - c2) v_I.v_J + (3/2 c0 - 2 c1 + c2/2) p_I p_J
// (D0 -> 3/2 p_I p_J, D1 -> 2 v_I.v_J - 2 p_I p_J, D2 -> d_IJ - 3/2 p_I p_J,
// and the reference's D2 is D2 - D1).  Pairs with one sphere inside the other
// (r <= |ai - aj|) take the reference's limiting forms.
RSD_HD void yamakawa_block(double x, double y, double z, double ai, double aj, double w, PairOut& o) {
  const double d = sqrt(x * x + y * y + z * z), id = 1.0 / d;
  const double t[3] = {x * id, y * id, z * id};
  const double d2 = d * d, d3 = d2 * d, d4 = d2 * d2, d5 = d4 * d;
  const bool inside = d <= fabs(ai - aj);
  const double amax = ai <= aj ? aj : ai;
  const double i2 = ai * ai, j2 = aj * aj, i3 = i2 * ai, j3 = j2 * aj, dm = ai - aj, dm2 = dm * dm;
  const double s4 = i2 + 4.0 * ai * aj + j2;
  // TT (:326-370) and RR (:373-438): delta and t t coefficients
  double tu, tr_, ru, rr_;
  if (inside) {
    tu = 1.0 / (6.0 * amax);
    ru = 1.0 / (8.0 * amax * amax * amax);
    tr_ = rr_ = 0.0;
  } else {
    const double pt = 1.0 / (192.0 * ai * aj), pr = 1.0 / (512.0 * i3 * j3);
    const double q = dm2 + 3.0 * d2, e = dm2 - d2;
    tu = pt * (16.0 * (ai + aj) - q * q / d3);
    tr_ = pt * 3.0 * e * e / d3;
    ru = pr * (5.0 * d3 - 27.0 * d * (i2 + j2) + 32.0 * (i3 + j3) - 9.0 * (i2 - j2) * (i2 - j2) / d -
               dm2 * dm2 * s4 / d3);
    rr_ = pr * 3.0 * e * e * (s4 - d2) / d3;
  }
  {
    const double tt[3] = {t[0], t[1], t[2]};
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ja[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) {
      const double dd = ia[q] == ja[q] ? 1.0 : 0.0, pp = tt[ia[q]] * tt[ja[q]];
      o.tt[q] += w * (tu * dd + tr_ * pp);
      o.rr[q] += w * (ru * dd + rr_ * pp);
    }
  }
  // TR (:467-505, taken for (-r, aj, ai) and transposed): TR(i,j) = c eps_ijk t_k
  {
    double c;
    if (inside)
      c = aj <= ai ? 0.0 : d / (8.0 * j3);
    else
      c = (dm - d) * (dm - d) * (i2 + 2.0 * ai * (aj + d) - 3.0 * (aj - d) * (aj - d)) / (128.0 * j3 * ai * d2);
    for (int m = 0; m < 3; ++m) o.tr[m] += w * c * t[m];
  }
  // TD (:582-640), RD (:512-556), DD (:665-765) coefficients
  double cU = 0.0, cR = 0.0, cH = 0.0, c0, c1, c2;
  if (inside) {
    if (ai <= aj) cU = -3.0 * d / (20.0 * j3);
    c0 = c1 = c2 = 3.0 / (20.0 * amax * amax * amax);
  } else {
    const double pre = 1.0 / (ai * j3), e = dm2 - d2;
    cU = pre * (10.0 * d2 - 24.0 * ai * d + 15.0 * (i2 - j2) - dm2 * dm2 * dm * (ai + 5.0 * aj) / d4) / 320.0;
    cR = pre * e * e * (dm * (ai + 5.0 * aj) - d2) / (128.0 * d4);
    cH = -3.0 / (256.0 * i3 * j3 * d3) * e * e * (s4 - d2);
    const double m6 = dm2 * dm2 * dm2 * (i2 + 6.0 * ai * aj + j2), m4 = dm2 * dm2 * s4;
    const double k = 3.0 / (1280.0 * i3 * j3);
    c0 = k * (3.0 * m6 / d5 - 10.0 * m4 / d3 + 32.0 * (i3 + j3) - 30.0 * (i2 + j2) * d + 5.0 * d3);
    c1 = k * (-2.0 * m6 / d5 + 5.0 * m4 / d3 - 15.0 * (i2 - j2) * (i2 - j2) / d + 32.0 * (i3 + j3) -
              25.0 * (i2 + j2) * d + 5.0 * d3);
    c2 = 0.5 * k * (m6 / d5 - 30.0 * (i2 - j2) * (i2 - j2) / d + 64.0 * (i3 + j3) - 40.0 * (i2 + j2) * d + 5.0 * d3);
  }
  YProj p;
  yproj(t[0], t[1], t[2], p);
  for (int I = 0; I < 5; ++I) {
    const double* v = p.Yx[I];
    const double vxt[3] = {v[1] * t[2] - v[2] * t[1], v[2] * t[0] - v[0] * t[2], v[0] * t[1] - v[1] * t[0]};
    for (int i = 0; i < 3; ++i) {
      o.gq[i + 3 * I] += w * (cU * v[i] + cR * t[i] * p.xYx[I]);
      o.hq[i + 3 * I] += w * cH * vxt[i];
    }
  }
  const double cy = 2.0 * (c1 - c2), cp = 1.5 * c0 - 2.0 * c1 + 0.5 * c2;
  int kq = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++kq) {
      const double vv = p.Yx[I][0] * p.Yx[J][0] + p.Yx[I][1] * p.Yx[J][1] + p.Yx[I][2] * p.Yx[J][2];
      o.qq[kq] += w * ((I == J ? c2 : 0.0) + cy * vv + cp * p.xYx[I] * p.xYx[J]);
    }
}

// The free-space RPY blocks of a pair (ROTNE_PRAGER_IJ_FTS, stokesian_Green.f90:11-108), weight w.
RSD_HD void free_rpy_block(double x, double y, double z, double ai, double aj, double w, PairOut& o) {
  Radial s;
  radial_oseen(sqrt(x * x + y * y + z * z), s);
  radial_block(x, y, z, s, ai * ai, aj * aj, 1.0, true, w, o);
  oseen_quirk(x, y, z, ai * ai, aj * aj, w, o);
}

// Overlap correction with correction_method = 1 (GRPY_YZD.f90:789-811):
// Yamakawa minus free-space RPY.
RSD_HD void overlap_yamakawa_correction(double x, double y, double z, double ai, double aj, PairOut& o) {
  yamakawa_block(x, y, z, ai, aj, 1.0, o);
  free_rpy_block(x, y, z, ai, aj, -1.0, o);
}

// ---------------------------------------------------------------- k space
// Per-wavevector weights (already divided by the cell volume LV), 64 doubles:
//   [0..5]  wTT = F/8            TT += wTT (1 - ct k^2) cos
//   [6..11] wRR = k^2 F/32       RR += wRR cos
//   [12..14] wTR = Phi k^2 k/16  tr += wTR (1 - ai^2 k^2/6) sin
//   [15..29] wG  = (F Y_I k)_i/8 GQ += wG (1 - cg k^2) sin
//   [30..44] wH  = Phi k^2 (k x Y_I k)_i/16   HQ += wH (1 - 0.1 aj^2 k^2) cos
//   [45..59] wQ  = Phi [k^2 (Y_I k . Y_J k) - (k Y_I k)(k Y_J k)]   QQ += wQ (1/8 - (ai^2+aj^2) k^2/80) cos
//   [60] k^2
// with F = Phi (delta k^2 - k k), Phi = 8 pi (1 + k^2/4l^2 + k^4/8l^4) exp(-k^2/4l^2) / k^4
// (Ewald_summation.f90:1057-1074 and the contractions of :340-412).
constexpr int kKW = 64;

RSD_HD void kweights(double kx, double ky, double kz, double lam, double inv_lv, double* w) {
  for (int i = 0; i < kKW; ++i) w[i] = 0.0;
  const double k2 = kx * kx + ky * ky + kz * kz;
  if (k2 == 0.0) return;
  const double k4 = k2 * k2, l2 = lam * lam;
  const double Phi = 8.0 * 3.141592653589793238462643383279502884 *
                     (1.0 + k2 / (4.0 * l2) + k4 / (8.0 * l2 * l2)) * exp(-k2 / (4.0 * l2)) / k4 * inv_lv;
  const double kv[3] = {kx, ky, kz};
  double F[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) F[i][j] = Phi * ((i == j ? k2 : 0.0) - kv[i] * kv[j]);
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    w[q] = F[si[q]][sj[q]] / 8.0;
    w[6 + q] = k2 * F[si[q]][sj[q]] / 32.0;
  }
  for (int m = 0; m < 3; ++m) w[12 + m] = Phi * k2 * kv[m] / 16.0;
  YProj p;
  yproj(kx, ky, kz, p);
  for (int I = 0; I < 5; ++I) {
    const double* v = p.Yx[I];
    const double cx[3] = {ky * v[2] - kz * v[1], kz * v[0] - kx * v[2], kx * v[1] - ky * v[0]};
    for (int i = 0; i < 3; ++i) {
      const double fyk = F[i][0] * v[0] + F[i][1] * v[1] + F[i][2] * v[2];
      w[15 + i + 3 * I] = fyk / 8.0;
      w[30 + i + 3 * I] = Phi * k2 * cx[i] / 16.0;
    }
  }
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k) {
      const double yy = p.Yx[I][0] * p.Yx[J][0] + p.Yx[I][1] * p.Yx[J][1] + p.Yx[I][2] * p.Yx[J][2];
      w[45 + k] = Phi * (k2 * yy - p.xYx[I] * p.xYx[J]);
    }
  w[60] = k2;
}

// Accumulate one wavevector for one pair: c = cos(k.R), s = sin(k.R).
RSD_HD void kspace_accumulate(const double* w, double c, double s, double ai2, double aj2, double phix, PairOut& o) {
  const double k2 = w[60];
  const double ft = c * (1.0 - (ai2 + aj2) / 6.0 * phix * k2);
  const double fr = s * (1.0 - ai2 * k2 / 6.0);
  const double fg = s * (1.0 - (ai2 / 6.0 + 0.1 * aj2) * k2);
  const double fh = c * (1.0 - 0.1 * aj2 * k2);
  const double fq = c * (0.125 - (ai2 + aj2) * k2 / 80.0);
  for (int q = 0; q < 6; ++q) {
    o.tt[q] += w[q] * ft;
    o.rr[q] += w[6 + q] * c;
  }
  for (int m = 0; m < 3; ++m) o.tr[m] += w[12 + m] * fr;
  for (int q = 0; q < 15; ++q) {
    o.gq[q] += w[15 + q] * fg;
    o.hq[q] += w[30 + q] * fh;
    o.qq[q] += w[45 + q] * fq;
  }
}

// Analytic self terms (Ewald_summation.f90:179-189), added after scaling:
// TT d_ij (1 - 6 l a/sqrt(pi) + 40/(3 sqrt(pi)) l^3 a^3 phix) / (6a), RR d_ij (3/4 - 10/sqrt(pi) a^3 l^3)/(6a^3),
// QQ d_IJ (9/10 - l^3 a^3/(25 sqrt(pi)) (50 - 126 l^2 a^2) 6)/(6a^3).
RSD_HD void self_analytic(double a, double lam, double phix, double& tt, double& rr, double& qq) {
  const double sp = 1.7724538509055160272981674833411452;
  const double l3a3 = lam * lam * lam * a * a * a;
  tt = (1.0 - 6.0 * lam / sp * a + 40.0 / 3.0 / sp * l3a3 * phix) / (6.0 * a);
  rr = (0.75 - 10.0 / sp * l3a3) / (6.0 * a * a * a);
  qq = (0.9 - l3a3 / 25.0 / sp * (50.0 - 126.0 * lam * lam * a * a) * 6.0) / (6.0 * a * a * a);
}

}  // namespace rsdm
