// Host build of the GPU pair-math header for CPU unit tests: the same
// closed-form expressions the HIP kernels evaluate, summed over the image
// cube here on the CPU so they can be compared with the literal oracle.
#include <cmath>
#include <cstring>

#include "../../regularized-stokesian-dynamics_amd/csrc/rsd_pairmath.h"

using namespace rsdm;

static void to_blocks(const PairOut& o, double* TT, double* RR, double* TR, double* QQ, double* GQ, double* HQ) {
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    TT[si[q] + 3 * sj[q]] = TT[sj[q] + 3 * si[q]] = o.tt[q];
    RR[si[q] + 3 * sj[q]] = RR[sj[q] + 3 * si[q]] = o.rr[q];
  }
  const double t0 = o.tr[0], t1 = o.tr[1], t2 = o.tr[2];
  double T[9] = {0, -t2, t1, t2, 0, -t0, -t1, t0, 0};  // TR(i,j) = eps_ijm t_m, column-major
  std::memcpy(TR, T, sizeof T);
  for (int q = 0; q < 15; ++q) {
    GQ[q] = o.gq[q];
    HQ[q] = o.hq[q];
  }
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k) QQ[I + 5 * J] = QQ[J + 5 * I] = o.qq[k];
}

extern "C" {

// PER_ROTNE_PRAGER_IJ equivalent: R already minimum-imaged.
void pm_rp_ij(const double* LR, const double* LI, double LV, double lam, double phi, int NR, int NI, const double* R,
              double ai, double aj, double* TT, double* RR, double* TR, double* QQ, double* GQ, double* HQ) {
  PairOut o;
  pairout_zero(o);
  const double phix = 1 - 0.2 * phi, ai2 = ai * ai, aj2 = aj * aj;
  Radial s;
  for (int nx = -NR; nx <= NR; ++nx)
    for (int ny = -NR; ny <= NR; ++ny)
      for (int nz = -NR; nz <= NR; ++nz) {
        double x[3];
        for (int i = 0; i < 3; ++i) x[i] = R[i] + LR[i] * nx + LR[i + 3] * ny + LR[i + 6] * nz;
        double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        radial_ewald(r, lam, s);
        radial_block(x[0], x[1], x[2], s, ai2, aj2, phix, true, 1.0, o);
      }
  double w[kKW];
  for (int nx = -NI; nx <= NI; ++nx)
    for (int ny = -NI; ny <= NI; ++ny)
      for (int nz = -NI; nz <= NI; ++nz) {
        if (!nx && !ny && !nz) continue;
        double kv[3];
        for (int j = 0; j < 3; ++j) kv[j] = 2 * M_PI * (nx * LI[0 + 3 * j] + ny * LI[1 + 3 * j] + nz * LI[2 + 3 * j]);
        kweights(kv[0], kv[1], kv[2], lam, 1.0 / LV, w);
        double kr = kv[0] * R[0] + kv[1] * R[1] + kv[2] * R[2];
        kspace_accumulate(w, std::cos(kr), std::sin(kr), ai2, aj2, phix, o);
      }
  to_blocks(o, TT, RR, TR, QQ, GQ, HQ);
}

void pm_rp_self(const double* LR, const double* LI, double LV, double lam, double phi, int NR, int NI, double a,
                double* TT, double* RR, double* QQ) {
  PairOut o;
  pairout_zero(o);
  const double phix = 1 - 0.2 * phi, a2 = a * a;
  Radial s;
  for (int nx = -NR; nx <= NR; ++nx)
    for (int ny = -NR; ny <= NR; ++ny)
      for (int nz = -NR; nz <= NR; ++nz) {
        if (!nx && !ny && !nz) continue;
        double x[3];
        for (int i = 0; i < 3; ++i) x[i] = LR[i] * nx + LR[i + 3] * ny + LR[i + 6] * nz;
        double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        radial_ewald(r, lam, s);
        radial_block(x[0], x[1], x[2], s, a2, a2, phix, false, 1.0, o);
      }
  double w[kKW];
  for (int nx = -NI; nx <= NI; ++nx)
    for (int ny = -NI; ny <= NI; ++ny)
      for (int nz = -NI; nz <= NI; ++nz) {
        if (!nx && !ny && !nz) continue;
        double kv[3];
        for (int j = 0; j < 3; ++j) kv[j] = 2 * M_PI * (nx * LI[0 + 3 * j] + ny * LI[1 + 3 * j] + nz * LI[2 + 3 * j]);
        kweights(kv[0], kv[1], kv[2], lam, 1.0 / LV, w);
        kspace_accumulate(w, 1.0, 0.0, a2, a2, phix, o);
      }
  double tt, rr, qq;
  self_analytic(a, lam, phix, tt, rr, qq);
  o.tt[0] += tt; o.tt[1] += tt; o.tt[2] += tt;
  o.rr[0] += rr; o.rr[1] += rr; o.rr[2] += rr;
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k)
      if (I == J) o.qq[k] += qq;
  double TR[9], GQ[15], HQ[15];
  to_blocks(o, TT, RR, TR, QQ, GQ, HQ);
}

void pm_overlap(const double* R, double ai, double aj, double* TT, double* RR, double* TR, double* QQ, double* GQ,
                double* HQ) {
  PairOut o;
  pairout_zero(o);
  overlap_regular_correction(R[0], R[1], R[2], ai, aj, o);
  to_blocks(o, TT, RR, TR, QQ, GQ, HQ);
}

void pm_yamakawa(const double* R, double ai, double aj, double* TT, double* RR, double* TR, double* QQ, double* GQ,
                 double* HQ) {
  PairOut o;
  pairout_zero(o);
  yamakawa_block(R[0], R[1], R[2], ai, aj, 1.0, o);
  to_blocks(o, TT, RR, TR, QQ, GQ, HQ);
}

void pm_overlap_yamakawa(const double* R, double ai, double aj, double* TT, double* RR, double* TR, double* QQ,
                         double* GQ, double* HQ) {
  PairOut o;
  pairout_zero(o);
  overlap_yamakawa_correction(R[0], R[1], R[2], ai, aj, o);
  to_blocks(o, TT, RR, TR, QQ, GQ, HQ);
}
}

#include "../../regularized-stokesian-dynamics_amd/csrc/rsd_lubmath.h"

extern "C" {

// lubcalc on tables given as the row-major contents of r2babc.dat (39x11),
// r2bgh.dat (47x7) and r2bm.dat (47x4); v[19] in lubcalc's order.
void pm_lubcalc(double r, const double* abc, const double* gh, const double* m, double* v) {
  static LubTables t;
  std::memcpy(t.abc, abc, sizeof t.abc);
  std::memcpy(t.gh, gh, sizeof t.gh);
  std::memcpy(t.m, m, sizeof t.m);
  lubcalc(r, t, v);
}
}
