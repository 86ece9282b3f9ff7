/*
 * rsd_oracle.c -- CPU restatement of the regularized Stokesian dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: it is loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, never by the product path.
 *
 * It restates, formula by formula, the FP64 algorithm of the reference
 * Fortran code (sherlockyzd/Regularized-Stokesian-Dynamics, src/), keeping
 * its array layouts (column-major, Fortran order) and its quirks.  Every
 * function cites the reference routine it follows.  The tensor algebra is
 * kept literal (full 3x3x3x3 arrays) on purpose: the GPU kernels use a
 * compact isotropic reformulation, and this literal form is what that
 * reformulation is checked against.
 *
 * Dense linear algebra (DPOTRF/DPOTRI) is not here: rsd_oracle.py calls
 * LAPACK through scipy for it, exactly as the reference calls LAPACK.
 *
 * Parity pins (see tests/): every function below is compared against the
 * flang-built reference (oracle/_ref) on seeded inputs and against the
 * committed golden fixtures in tests/golden/.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define T2(a, i, j) (a)[(i) + 3 * (j)]
#define T3(a, i, j, k) (a)[(i) + 3 * (j) + 9 * (k)]
#define T4(a, i, j, k, l) (a)[(i) + 3 * (j) + 9 * (k) + 27 * (l)]
#define Y2I(I, i, j) g_Y2[(I) + 5 * (i) + 15 * (j)]

static const double PAI = 3.141592653589793238462643383279502884;
static const double PIP5 = 0.564189583547756286948079451560772585844;

/* ---------------------------------------------------------------- tensors */
static int g_kmax2;  /* spherical reciprocal cutoff, see orc_set_images */
/* tensors.f90:2-14 INIT_KDirac, :20-37 INIT_EPS, :43-67 INIT_Y2, :132-172 INIT_II */
static double g_KD[9], g_EPS[27], g_Y2[45], g_II[81], g_IIiso[81];
static int g_init = 0;

void orc_init(void) {
  memset(g_KD, 0, sizeof g_KD);
  memset(g_EPS, 0, sizeof g_EPS);
  memset(g_Y2, 0, sizeof g_Y2);
  for (int i = 0; i < 3; ++i) T2(g_KD, i, i) = 1.0;
  T3(g_EPS, 0, 1, 2) = 1.0;
  T3(g_EPS, 0, 2, 1) = -1.0;
  T3(g_EPS, 1, 0, 2) = -1.0;
  T3(g_EPS, 1, 2, 0) = 1.0;
  T3(g_EPS, 2, 0, 1) = 1.0;
  T3(g_EPS, 2, 1, 0) = -1.0;
  const double s2 = 1.0 / sqrt(2.0), s6 = 1.0 / sqrt(6.0);
  Y2I(0, 0, 1) = s2;
  Y2I(0, 1, 0) = s2;
  Y2I(1, 1, 2) = -s2;
  Y2I(1, 2, 1) = -s2;
  Y2I(2, 0, 0) = -s6;
  Y2I(2, 1, 1) = -s6;
  Y2I(2, 2, 2) = sqrt(2.0 / 3.0);
  Y2I(3, 0, 2) = -s2;
  Y2I(3, 2, 0) = -s2;
  Y2I(4, 0, 0) = s2;
  Y2I(4, 1, 1) = -s2;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m) {
          T4(g_II, i, j, l, m) = (3.0 * T2(g_KD, i, m) * T2(g_KD, j, l) + 3.0 * T2(g_KD, i, l) * T2(g_KD, j, m) -
                                  2.0 * T2(g_KD, i, j) * T2(g_KD, l, m)) / 6.0;
          T4(g_IIiso, i, j, l, m) = T2(g_KD, i, m) * T2(g_KD, j, l) + T2(g_KD, i, l) * T2(g_KD, j, m) +
                                    T2(g_KD, i, j) * T2(g_KD, l, m);
        }
  g_init = 1;
}
static void ensure_init(void) {
  if (!g_init) orc_init();
}

/* tensors.f90:993-1008 QQ_TR (Y21 . B . Y22 with Y21=Y22=Y2) */
static void qq_tr(double B1[25], const double Bc[81]) {
  for (int I = 0; I < 5; ++I)
    for (int J = 0; J < 5; ++J) {
      double a = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) a += T4(Bc, i, j, k, l) * Y2I(I, i, j) * Y2I(J, k, l);
      B1[I + 5 * J] = a;
    }
}
/* tensors.f90:1010-1023 GH_TR -> RTD_back_transY2 :952-964 (contract last two indices) */
static void gh_tr(double C1[15], const double G[27]) {
  for (int I = 0; I < 5; ++I)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v += T3(G, c, i, j) * Y2I(I, i, j);
      C1[c + 3 * I] = v;
    }
}
/* tensors.f90:966-978 RTD_front_transY2 (contract first two indices) -> C1(5,3) */
static void rtd_front(double C1[15], const double G[27]) {
  for (int I = 0; I < 5; ++I)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v += T3(G, i, j, c) * Y2I(I, i, j);
      C1[I + 5 * c] = v;
    }
}
/* tensors.f90:890-903 EI_transY2 */
void orc_ei_transy2(double ei[5], const double E[9]) {
  ensure_init();
  for (int I = 0; I < 5; ++I) {
    double a = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) a += T2(E, i, j) * Y2I(I, i, j);
    ei[I] = a;
  }
}

/* ---------------------------------------------------------------- lattice */
typedef struct {
  double LR[9], LI[9], LV;
  double sig, lam, logerr, phi;
  int NR, NI;
} orc_lattice;

/* tensors.f90:704-726 MAT_INV3 */
static void mat_inv3(double LI[9], const double LR[9], double *det) {
#define L(i, j) LR[(i - 1) + 3 * (j - 1)]
#define O(i, j) LI[(i - 1) + 3 * (j - 1)]
  double d = L(1, 1) * L(2, 2) * L(3, 3) + L(1, 2) * L(2, 3) * L(3, 1) + L(1, 3) * L(2, 1) * L(3, 2) -
             L(1, 3) * L(2, 2) * L(3, 1) - L(2, 3) * L(3, 2) * L(1, 1) - L(3, 3) * L(1, 2) * L(2, 1);
  O(1, 1) = (1.0 / d) * (L(2, 2) * L(3, 3) - L(2, 3) * L(3, 2));
  O(1, 2) = (1.0 / d) * (L(1, 3) * L(3, 2) - L(1, 2) * L(3, 3));
  O(1, 3) = (1.0 / d) * (L(1, 2) * L(2, 3) - L(1, 3) * L(2, 2));
  O(2, 1) = (1.0 / d) * (L(2, 3) * L(3, 1) - L(2, 1) * L(3, 3));
  O(2, 2) = (1.0 / d) * (L(1, 1) * L(3, 3) - L(1, 3) * L(3, 1));
  O(2, 3) = (1.0 / d) * (L(1, 3) * L(2, 1) - L(1, 1) * L(2, 3));
  O(3, 1) = (1.0 / d) * (L(2, 1) * L(3, 2) - L(2, 2) * L(3, 1));
  O(3, 2) = (1.0 / d) * (L(1, 2) * L(3, 1) - L(1, 1) * L(3, 2));
  O(3, 3) = (1.0 / d) * (L(1, 1) * L(2, 2) - L(1, 2) * L(2, 1));
  *det = d;
#undef L
#undef O
}

/* csd_post.f90:105-136 SET_LATTICE_SKEW (steady shear branch) */
void orc_set_lattice_skew(double T, const double LB[3], double gamma, double LR[9], double LI[9], double *LV) {
  memset(LR, 0, 9 * sizeof(double));
  LR[0] = LB[0];
  LR[4] = LB[1];
  LR[8] = LB[2];
  if (gamma != 0.0) {
    double tshear = LB[0] / (LB[2] * gamma);
    double tmod = fmod(T, tshear); /* Fortran MOD: sign of the dividend */
    LR[0 + 3 * 2] = LB[2] * tmod * gamma;
  } else {
    LR[0 + 3 * 2] = 0.0;
  }
  double det;
  mat_inv3(LI, LR, &det);
  *LV = fabs(det);
}

/* Eigenvalues of a symmetric 3x3 matrix by cyclic Jacobi rotations.
 * Restates tensors.f90:679-698 EIGEN (DSYTRD/DORGTR/DSTEQR): only the
 * eigenvalues are used by the caller, and any convergent symmetric
 * eigensolver gives them to rounding. */
static void sym3_eig(const double A[9], double ev[3]) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = A[i + 3 * j];
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  ev[0] = a[0][0];
  ev[1] = a[1][1];
  ev[2] = a[2][2];
}

/* csd_post.f90:158-183 CALC_LATTICE_INVERSE.  phi uses the particle radii
 * (the reference reads CONFIG::RADII and SIZE::NN). */
void orc_calc_lattice_inverse(orc_lattice *lat, double ews_err, int n, const double *radii) {
  lat->sig = pow(lat->LV, 1.0 / 3.0) / sqrt(2.0 * PAI);
  lat->lam = sqrt(PAI) / pow(lat->LV, 1.0 / 3.0);
  lat->logerr = -log(ews_err);
  double MM[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += lat->LR[k + 3 * i] * lat->LR[k + 3 * j];
      MM[i + 3 * j] = s;
    }
  double ev[3];
  sym3_eig(MM, ev);
  double emin = fmin(ev[0], fmin(ev[1], ev[2])), emax = fmax(ev[0], fmax(ev[1], ev[2]));
  double rnr = 2.0 * lat->logerr * lat->sig * lat->sig / emin;
  lat->NR = (int)(sqrt(rnr) + 1.0);
  double lmin = 1.0 / emax;
  double rni = lat->logerr / (2.0 * PAI * PAI * lat->sig * lat->sig * lmin);
  lat->NI = (int)(sqrt(rni) + 1.0);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += radii[i] * radii[i] * radii[i];
  lat->phi = s * 4.0 / 3.0 * PAI / lat->LV;
}

/* csd_post.f90:24-63 PER_SKEW_CORRECTION */
void orc_per_skew_correction(const double LR[9], const double LI[9], double R[3], double *dmin) {
  double X[3], RC[3], RMIN[3];
  for (int i = 0; i < 3; ++i) X[i] = LI[i] * R[0] + LI[i + 3] * R[1] + LI[i + 6] * R[2];
  for (int i = 0; i < 3; ++i) {
    X[i] = fmod(X[i], 1.0);
    if (X[i] < 0.0) X[i] += 1.0;
  }
  for (int i = 0; i < 3; ++i) RC[i] = LR[i] * X[0] + LR[i + 3] * X[1] + LR[i + 6] * X[2];
  double D = RC[0] * RC[0] + RC[1] * RC[1] + RC[2] * RC[2];
  for (int i = 0; i < 3; ++i) RMIN[i] = RC[i];
  for (int m1 = 0; m1 <= 1; ++m1)
    for (int m2 = 0; m2 <= 1; ++m2)
      for (int m3 = 0; m3 <= 1; ++m3) {
        if (m1 == 0 && m2 == 0 && m3 == 0) continue;
        double RL[3], d = 0.0;
        for (int i = 0; i < 3; ++i) RL[i] = LR[i] * m1 + LR[i + 3] * m2 + LR[i + 6] * m3;
        for (int i = 0; i < 3; ++i) d += (RC[i] - RL[i]) * (RC[i] - RL[i]);
        if (d < D) {
          D = d;
          for (int i = 0; i < 3; ++i) RMIN[i] = RC[i] - RL[i];
        }
      }
  for (int i = 0; i < 3; ++i) R[i] = RMIN[i];
  *dmin = sqrt(D);
}

/* csd_post.f90:65-102 PERIOD_CORRECTION (steady shear) */
void orc_period_correction(double R[3], const double LB[3], double T, double gamma) {
  double lsh = LB[2] * gamma * T;
  if (R[2] > LB[2]) {
    R[2] -= LB[2];
    R[0] = fmod(R[0] - lsh, LB[0]);
  }
  if (R[2] < 0.0) {
    R[2] += LB[2];
    R[0] = fmod(R[0] + lsh, LB[0]);
  }
  if (R[0] > LB[0])
    R[0] -= LB[0];
  else if (R[0] < 0.0)
    R[0] += LB[0];
  if (R[1] > LB[1])
    R[1] -= LB[1];
  else if (R[1] < 0.0)
    R[1] += LB[1];
}

/* ----------------------------------------------- literal tensor helpers */
static void kijl(const double *rn, double *k) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l)
        T3(k, i, j, l) = T2(g_KD, i, j) * rn[l] + T2(g_KD, i, l) * rn[j] + T2(g_KD, j, l) * rn[i];
}
static void kijlm(const double *rn, double *k) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m)
          T4(k, i, j, l, m) = T2(g_KD, i, j) * rn[l] * rn[m] + T2(g_KD, i, l) * rn[j] * rn[m] +
                              T2(g_KD, j, l) * rn[i] * rn[m] + T2(g_KD, i, m) * rn[j] * rn[l] +
                              T2(g_KD, j, m) * rn[i] * rn[l] + T2(g_KD, l, m) * rn[i] * rn[j];
}
static void rijl(const double *rn, double *r) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l) T3(r, i, j, l) = rn[l] * rn[j] * rn[i];
}
static void rijlm(const double *rn, double *r) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m) T4(r, i, j, l, m) = rn[l] * rn[j] * rn[i] * rn[m];
}

/* The radial-function derivative tensors a kernel family provides.  Both
 * the Ewald r*erfc(lr) family (Ewald_summation.f90:658-1053) and the
 * regularised family (Regularization.f90:400-778) fill this struct; the
 * downstream J/R/K algebra is shared (Ewald_summation.f90:427-655 and
 * Regularization.f90:169-392 are the same formulas). */
typedef struct {
  double DD[9], D2, D4, DDD[27], DDDD[81], DD2[3], DDD2[9], DDDD2[27], DDDDD2[81], DD4[3], DDD4[9];
} radial_t;

/* Ewald_summation.f90:658-743 (E_def, erfc1..6) and :805-1053 (derivative tensors) */
static void rerfc_tensors(const double *rn, double lam, radial_t *t) {
  double r2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
  double r = sqrt(r2);
  double rlam2 = r * r * lam * lam, rlam4 = rlam2 * rlam2;
  double E = 2.0 * PIP5 * lam * exp(-rlam2);
  double e1 = -E;
  double e2 = 2.0 * lam * lam * r * E;
  double e3 = -2.0 * lam * lam * E * (2.0 * rlam2 - 1.0);
  double e4 = 4.0 * lam * lam * lam * lam * r * E * (2.0 * rlam2 - 3.0);
  double e5 = -4.0 * lam * lam * lam * lam * E * (4.0 * rlam4 - 12.0 * rlam2 + 3.0);
  double e6 = 8.0 * lam * lam * lam * lam * lam * lam * r * E * (4.0 * rlam4 - 20.0 * rlam2 + 15.0);
  double ec = erfc(lam * r);
  double r3 = r * r * r, r4 = r2 * r2, r5 = r3 * r2, r6 = r5 * r, r7 = r6 * r, r8 = r7 * r, r9 = r8 * r;
  double kl[27], rl[27], klm[81], rlm[81];
  kijl(rn, kl);
  rijl(rn, rl);
  kijlm(rn, klm);
  rijlm(rn, rlm);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double kd = T2(g_KD, i, j), xx = rn[i] * rn[j];
      T2(t->DD, i, j) = (kd / r - xx / r3) * ec + (kd + xx / r2) * e1 + xx / r * e2;
      T2(t->DDD2, i, j) = (-2.0 * kd / r3 + 6.0 * xx / r5) * ec - (-2.0 * kd / r2 + 6.0 * xx / r4) * e1 +
                          (5.0 * kd / r - 3.0 * xx / r3) * e2 + (kd + 5.0 * xx / r2) * e3 + xx / r * e4;
      T2(t->DDD4, i, j) = (-12.0 * kd / r3 + 36.0 * xx / r5) * e2 - (-12.0 * kd / r2 + 36.0 * xx / r4) * e3 +
                          (9.0 * kd / r + 3.0 * xx / r3) * e4 + (kd + 9.0 * xx / r2) * e5 + xx / r * e6;
    }
  t->D2 = 2.0 / r * ec + 4.0 * e1 + r * e2;
  t->D4 = 12.0 / r * e2 + 8.0 * e3 + r * e4;
  for (int l = 0; l < 3; ++l) {
    t->DD2[l] = -2.0 * rn[l] / r3 * ec + 2.0 * rn[l] / r2 * e1 + 5.0 * rn[l] / r * e2 + rn[l] * e3;
    t->DD4[l] = -12.0 * rn[l] / r3 * e2 + 12.0 * rn[l] / r2 * e3 + 9.0 * rn[l] / r * e4 + rn[l] * e5;
  }
  for (int q = 0; q < 27; ++q) {
    t->DDD[q] = (-kl[q] / r3 + 3.0 * rl[q] / r5) * ec - (-kl[q] / r2 + 3.0 * rl[q] / r4) * e1 + kl[q] / r * e2 +
                rl[q] / r2 * e3;
    t->DDDD2[q] = (6.0 * kl[q] / r5 - 30.0 * rl[q] / r7) * ec - (6.0 * kl[q] / r4 - 30.0 * rl[q] / r6) * e1 +
                  (-3.0 * kl[q] / r3 + 3.0 * rl[q] / r5) * e2 + (5.0 * kl[q] / r2 - 13.0 * rl[q] / r4) * e3 +
                  (kl[q] / r + 4.0 * rl[q] / r3) * e4 + rl[q] / r2 * e5;
  }
  for (int q = 0; q < 81; ++q) {
    double I = g_IIiso[q], k = klm[q], x = rlm[q];
    t->DDDD[q] = (-I / r3 + 3.0 * k / r5 - 15.0 * x / r7) * ec - (-I / r2 + 3.0 * k / r4 - 15.0 * x / r6) * e1 +
                 (I / r - 3.0 * x / r5) * e2 + (k / r2 - 2.0 * x / r4) * e3 + x / r3 * e4;
    t->DDDDD2[q] = (6.0 * I / r5 - 30.0 * k / r7 + 210.0 * x / r9) * ec -
                   (6.0 * I / r4 - 30.0 * k / r6 + 210.0 * x / r8) * e1 +
                   (-3.0 * I / r3 + 3.0 * k / r5 + 15.0 * x / r7) * e2 +
                   (5.0 * I / r2 - 13.0 * k / r4 + 55.0 * x / r6) * e3 + (I / r + 4.0 * k / r3 - 25.0 * x / r5) * e4 +
                   (k / r2 + 2.0 * x / r4) * e5 + x / r3 * e6;
  }
}

/* Regularization.f90:400-472 (E_def, erf1..5) and :538-778 (phir tensors) */
static void phir_tensors(const double *rn, double kapa, radial_t *t) {
  double r2 = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
  double r = sqrt(r2);
  double rlam2 = r * r * kapa * kapa, rlam4 = rlam2 * rlam2;
  double E = 2.0 * PIP5 * kapa * exp(-rlam2);
  double e1 = E;
  double e2 = -2.0 * kapa * kapa * r * E;
  double e3 = 2.0 * kapa * kapa * E * (2.0 * rlam2 - 1.0);
  double e4 = -4.0 * kapa * kapa * kapa * kapa * r * E * (2.0 * rlam2 - 3.0);
  double e5 = 4.0 * kapa * kapa * kapa * kapa * E * (4.0 * rlam4 - 12.0 * rlam2 + 3.0);
  double ef = erf(kapa * r);
  double r3 = r * r * r, r4 = r2 * r2, r5 = r3 * r2, r6 = r5 * r, r7 = r6 * r, r8 = r7 * r, r9 = r8 * r;
  double kl[27], rl[27], klm[81], rlm[81];
  kijl(rn, kl);
  rijl(rn, rl);
  kijlm(rn, klm);
  rijlm(rn, rlm);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double kd = T2(g_KD, i, j), xx = rn[i] * rn[j];
      T2(t->DD, i, j) = (kd / r - xx / r3) * ef + xx / r2 * e1;
      T2(t->DDD2, i, j) = (-2.0 * kd / r3 + 6.0 * xx / r5) * ef + (2.0 * kd / r2 - 6.0 * xx / r4) * e1 +
                          (kd / r + xx / r3) * e2 + xx / r2 * e3;
      T2(t->DDD4, i, j) = (-4.0 * kd / r3 + 12.0 * xx / r5) * e2 + (4.0 * kd / r2 - 12.0 * xx / r4) * e3 +
                          (kd / r + 3.0 * xx / r3) * e4 + xx / r2 * e5;
    }
  t->D2 = 2.0 / r * ef + e1;
  t->D4 = 4.0 / r * e2 + e3;
  for (int l = 0; l < 3; ++l) {
    t->DD2[l] = -2.0 * rn[l] / r3 * ef + 2.0 * rn[l] / r2 * e1 + rn[l] / r * e2;
    t->DD4[l] = -4.0 * rn[l] / r3 * e2 + 4.0 * rn[l] / r2 * e3 + rn[l] / r * e4;
  }
  for (int q = 0; q < 27; ++q) {
    t->DDD[q] = (-kl[q] / r3 + 3.0 * rl[q] / r5) * ef + (kl[q] / r2 - 3.0 * rl[q] / r4) * e1 + rl[q] / r3 * e2;
    t->DDDD2[q] = (6.0 * kl[q] / r5 - 30.0 * rl[q] / r7) * ef + (-6.0 * kl[q] / r4 + 30.0 * rl[q] / r6) * e1 +
                  (kl[q] / r3 - 9.0 * rl[q] / r5) * e2 + (kl[q] / r2 - rl[q] / r4) * e3 + rl[q] / r3 * e4;
  }
  for (int q = 0; q < 81; ++q) {
    double I = g_IIiso[q], k = klm[q], x = rlm[q];
    t->DDDD[q] = (-I / r3 + 3.0 * k / r5 - 15.0 * x / r7) * ef + (I / r2 - 3.0 * k / r4 + 15.0 * x / r6) * e1 +
                 (k / r3 - 6.0 * x / r5) * e2 + x / r4 * e3;
    t->DDDDD2[q] = (6.0 * I / r5 - 30.0 * k / r7 + 210.0 * x / r9) * ef +
                   (-6.0 * I / r4 + 30.0 * k / r6 - 210.0 * x / r8) * e1 +
                   (I / r3 - 9.0 * k / r5 + 75.0 * x / r7) * e2 + (I / r2 - k / r4 - 5.0 * x / r6) * e3 +
                   (k / r3 - 4.0 * x / r5) * e4 + x / r4 * e5;
  }
}

/* J, R (rotlet), K (stresslet) tensors and derivatives:
 * Ewald_summation.f90:427-655 Init_Jr .. Init_DD2_Kr */
typedef struct {
  double Jr[9], D_Jr[27], DD_Jr[81], D2_Jr[9], DD2_Jr[27], DDD2_Jr[81];
  double Ror[9], D_Ror[27], D2_Ror[9];
  double Kr[27], D_Kr[81], D2_Kr[27], DD2_Kr[81];
} jrk_t;

static void jrk_from_radial(const radial_t *t, jrk_t *o) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double kd = T2(g_KD, i, j);
      T2(o->Jr, i, j) = kd * t->D2 - T2(t->DD, i, j);
      T2(o->D2_Jr, i, j) = kd * t->D4 - T2(t->DDD2, i, j);
      for (int l = 0; l < 3; ++l) {
        T3(o->D_Jr, i, j, l) = kd * t->DD2[l] - T3(t->DDD, l, i, j);
        T3(o->DD2_Jr, i, j, l) = kd * t->DD4[l] - T3(t->DDDD2, l, i, j);
        for (int m = 0; m < 3; ++m) {
          T4(o->DD_Jr, i, j, m, l) = kd * T2(t->DDD2, m, l) - T4(t->DDDD, m, l, i, j);
          T4(o->DDD2_Jr, i, j, l, m) = kd * T2(t->DDD4, l, m) - T4(t->DDDDD2, l, m, i, j);
        }
      }
    }
  memset(o->Ror, 0, sizeof o->Ror);
  memset(o->D_Ror, 0, sizeof o->D_Ror);
  memset(o->D2_Ror, 0, sizeof o->D2_Ror);
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double e = T3(g_EPS, j, k, l);
          T2(o->Ror, i, j) -= 0.5 * e * T3(o->D_Jr, i, l, k);
          T2(o->D2_Ror, i, j) -= 0.5 * e * T3(o->DD2_Jr, i, l, k);
          for (int q = 0; q < 3; ++q) T3(o->D_Ror, i, j, q) -= 0.5 * e * T4(o->DD_Jr, i, l, q, k);
        }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        T3(o->Kr, i, j, k) = 0.5 * (T3(o->D_Jr, i, j, k) + T3(o->D_Jr, i, k, j));
        T3(o->D2_Kr, i, j, k) = 0.5 * (T3(o->DD2_Jr, i, j, k) + T3(o->DD2_Jr, i, k, j));
        for (int l = 0; l < 3; ++l) {
          T4(o->D_Kr, i, j, k, l) = 0.5 * (T4(o->DD_Jr, i, j, l, k) + T4(o->DD_Jr, i, k, l, j));
          T4(o->DD2_Kr, i, j, k, l) = 0.5 * (T4(o->DDD2_Jr, i, j, l, k) + T4(o->DDD2_Jr, i, k, l, j));
        }
      }
}

/* Accumulators for one pair (Cartesian, before projection) */
typedef struct {
  double TT[9], RR[9], TR[9], QQ[81], GQ[27], HQ[27];
} pairacc_t;

/* The per-image real-space accumulation of PER_ROTNE_PRAGER_IJ
 * (Ewald_summation.f90:302-334); also the body of Regular_ROTNE_PRAGER_IJ
 * (Regularization.f90:120-151) and ROTNE_PRAGER_IJ_FTS
 * (stokesian_Green.f90:64-100) with phix=1. */
static void accumulate_real(pairacc_t *A, const jrk_t *o, double ai, double aj, double phix, int with_tr_gh) {
  for (int q = 0; q < 9; ++q) {
    A->TT[q] += 0.75 * ai * o->Jr[q] + 0.75 * ai * (ai * ai + aj * aj) / 6.0 * o->D2_Jr[q] * phix;
  }
  double RRAA[9] = {0};
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T2(RRAA, i, j) += T3(g_EPS, i, k, l) * T3(o->D_Ror, l, j, k);
  for (int q = 0; q < 9; ++q) A->RR[q] += 3.0 * ai * ai * ai / 8.0 * RRAA[q];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double qa = 0.5 * (T4(o->D_Kr, i, k, l, j) + T4(o->D_Kr, j, k, l, i));
          double qb = 0.5 * (T4(o->DD2_Kr, i, k, l, j) + T4(o->DD2_Kr, j, k, l, i));
          T4(A->QQ, i, j, k, l) += -3.0 / 4.0 * ai * ai * ai * qa - 3.0 / 40.0 * ai * ai * ai * (ai * ai + aj * aj) * qb;
        }
  if (!with_tr_gh) return;
  for (int q = 0; q < 27; ++q) {
    double g = o->Kr[q] + (ai * ai / 6.0 + 0.1 * aj * aj) * o->D2_Kr[q];
    A->GQ[q] -= 0.75 * ai * g;
  }
  double HAA[27] = {0};
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            T3(HAA, i, j, k) += T3(g_EPS, i, l, m) * (T4(o->D_Kr, m, j, k, l) + 0.1 * aj * aj * T4(o->DD2_Kr, m, j, k, l));
  for (int q = 0; q < 27; ++q) A->HQ[q] -= 3.0 / 8.0 * ai * ai * ai * HAA[q];
  for (int q = 0; q < 9; ++q) A->TR[q] += ai * (3.0 / 4.0) * o->Ror[q] + ai * ai * ai / 8.0 * o->D2_Ror[q];
}

/* k-space tensors: Ewald_summation.f90:1057-1257 Init_F_* */
typedef struct {
  double Jk[9], DJk[27], DDJk[81], D2Jk[9], DD2Jk[27];
  double Rk[9], DRk[27], D2Rk[9], Kk[27], DKk[81], D2Kk[27], DD2Kk[81];
} ktens_t;

static void k_tensors(const double *kn, double lam, ktens_t *t) {
  double k2 = kn[0] * kn[0] + kn[1] * kn[1] + kn[2] * kn[2];
  double dist = sqrt(k2);
  k2 = dist * dist;
  double k4 = k2 * k2;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      T2(t->Jk, i, j) = (T2(g_KD, i, j) * k2 - kn[i] * kn[j]) / k4 * 8.0 * PAI *
                        (1 + k2 / (4.0 * lam * lam) + k4 / (8.0 * lam * lam * lam * lam)) * exp(-k2 / (4.0 * lam * lam));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T2(t->D2Jk, i, j) = -dist * dist * T2(t->Jk, i, j);
      for (int l = 0; l < 3; ++l) {
        T3(t->DJk, i, j, l) = kn[l] * T2(t->Jk, i, j);
        for (int m = 0; m < 3; ++m) T4(t->DDJk, i, j, m, l) = -kn[m] * kn[l] * T2(t->Jk, i, j);
      }
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l) T3(t->DD2Jk, i, j, l) = kn[l] * T2(t->D2Jk, i, j);
  memset(t->Rk, 0, sizeof t->Rk);
  memset(t->DRk, 0, sizeof t->DRk);
  memset(t->D2Rk, 0, sizeof t->D2Rk);
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double e = T3(g_EPS, j, m, n);
          T2(t->Rk, i, j) -= 0.5 * e * T3(t->DJk, i, n, m);
          T2(t->D2Rk, i, j) -= 0.5 * e * T3(t->DD2Jk, i, n, m);
          for (int l = 0; l < 3; ++l) T3(t->DRk, i, j, l) -= 0.5 * e * T4(t->DDJk, i, n, l, m);
        }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        T3(t->Kk, i, j, k) = 0.5 * (T3(t->DJk, i, j, k) + T3(t->DJk, i, k, j));
        T3(t->D2Kk, i, j, k) = T3(t->DD2Jk, i, j, k);
        for (int l = 0; l < 3; ++l) {
          T4(t->DKk, i, j, k, l) = 0.5 * (T4(t->DDJk, i, j, l, k) + T4(t->DDJk, i, k, l, j));
          T4(t->DD2Kk, i, j, k, l) = kn[k] * kn[l] * dist * dist * T2(t->Jk, i, j);
        }
      }
}

/* ------------------------------------------------ Ewald RPY self and pair */
/* Ewald_summation.f90:19-191 PER_ROTNE_PRAGER_SELF */
void orc_rp_self(const orc_lattice *lat, double aI, double TTS[9], double RRS[9], double QQS_TR[25]) {
  ensure_init();
  double phix = 1 - 0.2 * lat->phi, aJ = aI;
  double nondim_a = 1.0 / (6.0 * aI), nondim_a3 = 1.0 / (6.0 * aI * aI * aI);
  pairacc_t A, B;
  memset(&A, 0, sizeof A);
  memset(&B, 0, sizeof B);
  radial_t rt;
  jrk_t o;
  int NR = lat->NR, NI = lat->NI;
  for (int nx = -NR; nx <= NR; ++nx)
    for (int ny = -NR; ny <= NR; ++ny)
      for (int nz = -NR; nz <= NR; ++nz) {
        if (nx == 0 && ny == 0 && nz == 0) continue;
        double NL[3] = {nx, ny, nz}, RN[3];
        for (int i = 0; i < 3; ++i) RN[i] = lat->LR[i] * NL[0] + lat->LR[i + 3] * NL[1] + lat->LR[i + 6] * NL[2];
        rerfc_tensors(RN, lat->lam, &rt);
        jrk_from_radial(&rt, &o);
        accumulate_real(&A, &o, aI, aJ, phix, 0);
      }
  ktens_t kt;
  for (int nx = -NI; nx <= NI; ++nx)
    for (int ny = -NI; ny <= NI; ++ny)
      for (int nz = -NI; nz <= NI; ++nz) {
        if (nx == 0 && ny == 0 && nz == 0) continue;
        if (g_kmax2 > 0 && nx * nx + ny * ny + nz * nz > g_kmax2) continue;
        double NL[3] = {nx, ny, nz}, KN[3];
        for (int j = 0; j < 3; ++j)
          KN[j] = 2 * PAI * (NL[0] * lat->LI[0 + 3 * j] + NL[1] * lat->LI[1 + 3 * j] + NL[2] * lat->LI[2 + 3 * j]);
        k_tensors(KN, lat->lam, &kt);
        for (int q = 0; q < 9; ++q)
          B.TT[q] += 0.75 * aI * kt.Jk[q] + 1.0 / 8.0 * aI * (aI * aI + aJ * aJ) * phix * kt.D2Jk[q];
        double RRBA[9] = {0};
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) T2(RRBA, i, j) += T3(g_EPS, i, k, l) * T3(kt.DRk, l, j, k);
        for (int q = 0; q < 9; ++q) B.RR[q] += 3.0 / 8.0 * aI * aI * aI * RRBA[q];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
              for (int l = 0; l < 3; ++l) {
                double qa = 0.5 * (T4(kt.DKk, i, k, l, j) + T4(kt.DKk, j, k, l, i));
                double qb = 0.5 * (T4(kt.DD2Kk, i, k, l, j) + T4(kt.DD2Kk, j, k, l, i));
                T4(B.QQ, i, j, k, l) += -3.0 / 4.0 * aI * aI * aI * qa - 3.0 / 40.0 * aI * aI * aI * (aI * aI + aJ * aJ) * qb;
              }
      }
  double lam = lat->lam, LV = lat->LV, sp = sqrt(PAI);
  double ttsr = 1.0 - 6.0 * lam / sp * aI + 40.0 / 3.0 / sp * lam * lam * lam * aI * aI * aI * phix;
  double rrsr = 0.75 * (1.0 - 0.0 * lat->phi) - 10.0 / sp * aI * aI * aI * lam * lam * lam;
  double qqsr = 0.9 - lam * lam * lam * aI * aI * aI / 25.0 / sp * (50.0 - 126.0 * lam * lam * aI * aI) * 6.0;
  for (int q = 0; q < 9; ++q) {
    TTS[q] = (A.TT[q] + B.TT[q] / LV + ttsr * g_KD[q]) * nondim_a;
    RRS[q] = (rrsr * g_KD[q] + A.RR[q] + B.RR[q] / LV) * nondim_a3;
  }
  double QQ[81];
  for (int q = 0; q < 81; ++q) QQ[q] = (A.QQ[q] + B.QQ[q] / LV + qqsr * g_II[q]) * nondim_a3;
  qq_tr(QQS_TR, QQ);
}

/* Ewald_summation.f90:196-423 PER_ROTNE_PRAGER_IJ.  Outputs in Fortran order:
 * TTP,RRP,TRP (3,3), QQP_TR (5,5), GPQ_TR,HPQ_TR (3,5). */
void orc_rp_ij(const orc_lattice *lat, const double R[3], double aI, double aJ, double TTP[9], double RRP[9],
               double TRP[9], double QQP[25], double GPQ[15], double HPQ[15]) {
  ensure_init();
  double phix = 1 - 0.2 * lat->phi;
  double nondim_a = 1.0 / (6.0 * aI), nondim_a3 = 1.0 / (6.0 * aI * aI * aI);
  pairacc_t A, B;
  memset(&A, 0, sizeof A);
  memset(&B, 0, sizeof B);
  radial_t rt;
  jrk_t o;
  int NR = lat->NR, NI = lat->NI;
  for (int nx = -NR; nx <= NR; ++nx)
    for (int ny = -NR; ny <= NR; ++ny)
      for (int nz = -NR; nz <= NR; ++nz) {
        double NL[3] = {nx, ny, nz}, RN[3];
        for (int i = 0; i < 3; ++i) RN[i] = R[i] + (lat->LR[i] * NL[0] + lat->LR[i + 3] * NL[1] + lat->LR[i + 6] * NL[2]);
        rerfc_tensors(RN, lat->lam, &rt);
        jrk_from_radial(&rt, &o);
        accumulate_real(&A, &o, aI, aJ, phix, 1);
      }
  ktens_t kt;
  for (int nx = -NI; nx <= NI; ++nx)
    for (int ny = -NI; ny <= NI; ++ny)
      for (int nz = -NI; nz <= NI; ++nz) {
        if (nx == 0 && ny == 0 && nz == 0) continue;
        if (g_kmax2 > 0 && nx * nx + ny * ny + nz * nz > g_kmax2) continue;
        double NL[3] = {nx, ny, nz}, KN[3];
        for (int j = 0; j < 3; ++j)
          KN[j] = 2 * PAI * (NL[0] * lat->LI[0 + 3 * j] + NL[1] * lat->LI[1 + 3 * j] + NL[2] * lat->LI[2 + 3 * j]);
        k_tensors(KN, lat->lam, &kt);
        double kr = KN[0] * R[0] + KN[1] * R[1] + KN[2] * R[2];
        double c = cos(kr), s = sin(kr);
        for (int q = 0; q < 9; ++q)
          B.TT[q] += (0.75 * aI * kt.Jk[q] + 0.75 * aI * (aI * aI + aJ * aJ) / 6.0 * kt.D2Jk[q] * phix) * c;
        double RRBA[9] = {0};
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) T2(RRBA, i, j) += T3(g_EPS, i, k, l) * T3(kt.DRk, l, j, k);
        for (int q = 0; q < 9; ++q) B.RR[q] += 3.0 / 8.0 * aI * aI * aI * RRBA[q] * c;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
              for (int l = 0; l < 3; ++l) {
                double qa = 0.5 * (T4(kt.DKk, i, k, l, j) + T4(kt.DKk, j, k, l, i));
                double qb = 0.5 * (T4(kt.DD2Kk, i, k, l, j) + T4(kt.DD2Kk, j, k, l, i));
                T4(B.QQ, i, j, k, l) +=
                    (-3.0 / 4.0 * aI * aI * aI * qa - 3.0 / 40.0 * aI * aI * aI * (aI * aI + aJ * aJ) * qb) * c;
              }
        for (int q = 0; q < 27; ++q) {
          double g = kt.Kk[q] + (aI * aI / 6.0 + 0.1 * aJ * aJ) * kt.D2Kk[q];
          B.GQ[q] += 0.75 * aI * g * s;
        }
        double HBA[27] = {0};
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m)
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k)
                  T3(HBA, i, j, k) += T3(g_EPS, i, l, m) * (T4(kt.DKk, m, j, k, l) + 0.1 * aJ * aJ * T4(kt.DD2Kk, m, j, k, l));
        for (int q = 0; q < 27; ++q) B.HQ[q] -= 3.0 / 8.0 * aI * aI * aI * HBA[q] * c;
        for (int q = 0; q < 9; ++q) B.TR[q] -= (aI * (3.0 / 4.0) * kt.Rk[q] + aI * aI * aI / 8.0 * kt.D2Rk[q]) * s;
      }
  double LV = lat->LV, QQ[81], GQ[27], HQ[27];
  for (int q = 0; q < 9; ++q) {
    TTP[q] = (A.TT[q] + B.TT[q] / LV) * nondim_a;
    RRP[q] = (A.RR[q] + B.RR[q] / LV) * nondim_a3;
    TRP[q] = (A.TR[q] + B.TR[q] / LV) * nondim_a;
  }
  for (int q = 0; q < 81; ++q) QQ[q] = (A.QQ[q] + B.QQ[q] / LV) * nondim_a3;
  for (int q = 0; q < 27; ++q) {
    GQ[q] = (A.GQ[q] + B.GQ[q] / LV) * nondim_a;
    HQ[q] = (A.HQ[q] + B.HQ[q] / LV) * nondim_a3;
  }
  gh_tr(GPQ, GQ);
  gh_tr(HPQ, HQ);
  qq_tr(QQP, QQ);
}

/* --------------------------------------------- overlap corrections */
static void finish_pair(const pairacc_t *A, double aI, double TTP[9], double RRP[9], double TRP[9], double QQP[25],
                        double GPQ[15], double HPQ[15]) {
  double nondim_a = 1.0 / (6.0 * aI), nondim_a3 = 1.0 / (6.0 * aI * aI * aI);
  double QQ[81], GQ[27], HQ[27];
  for (int q = 0; q < 9; ++q) {
    TTP[q] = A->TT[q] * nondim_a;
    RRP[q] = A->RR[q] * nondim_a3;
    TRP[q] = A->TR[q] * nondim_a;
  }
  for (int q = 0; q < 81; ++q) QQ[q] = A->QQ[q] * nondim_a3;
  for (int q = 0; q < 27; ++q) {
    GQ[q] = A->GQ[q] * nondim_a;
    HQ[q] = A->HQ[q] * nondim_a3;
  }
  gh_tr(GPQ, GQ);
  gh_tr(HPQ, HQ);
  qq_tr(QQP, QQ);
}

/* Regularization.f90:18-162 Regular_ROTNE_PRAGER_IJ */
void orc_rp_regular(const double RN[3], double aI, double aJ, double TTP[9], double RRP[9], double TRP[9],
                    double QQP[25], double GPQ[15], double HPQ[15]) {
  ensure_init();
  double dist = sqrt(RN[0] * RN[0] + RN[1] * RN[1] + RN[2] * RN[2]);
  double kapa = sqrt(PAI) / 3.0 / fabs((aI + aJ) - dist);
  radial_t rt;
  jrk_t o;
  phir_tensors(RN, kapa, &rt);
  jrk_from_radial(&rt, &o);
  pairacc_t A;
  memset(&A, 0, sizeof A);
  accumulate_real(&A, &o, aI, aJ, 1.0, 1);
  finish_pair(&A, aI, TTP, RRP, TRP, QQP, GPQ, HPQ);
}

/* stokesian_Green.f90:11-108 ROTNE_PRAGER_IJ_FTS with its Init_*g helpers
 * (:215-434), kept literal: Init_DDD2_Jg's delta_ij x_l x_k term carries
 * the reference's sign, and D2_Kg/DD2_Kg are the unsymmetrised copies. */
void orc_rp_free_fts(const double RN[3], double aI, double aJ, double TTP[9], double RRP[9], double TRP[9],
                     double QQP[25], double GPQ[15], double HPQ[15]) {
  ensure_init();
  const double *rn = RN;
  double dist = sqrt(RN[0] * RN[0] + RN[1] * RN[1] + RN[2] * RN[2]);
  double d3 = pow(dist, 3.0), d5 = pow(dist, 5.0), d7 = pow(dist, 7.0), d9 = pow(dist, 9.0);
  jrk_t o;
  memset(&o, 0, sizeof o);
#define KD(i, j) T2(g_KD, i, j)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T2(o.Jr, i, j) = KD(i, j) / dist + rn[i] * rn[j] / d3;
      T2(o.D2_Jr, i, j) = 2.0 * KD(i, j) / d3 - 6.0 * rn[i] * rn[j] / d5;
      for (int l = 0; l < 3; ++l) {
        T3(o.D_Jr, i, j, l) = (-KD(i, j) * rn[l] + KD(i, l) * rn[j] + KD(j, l) * rn[i]) / d3 - 3.0 * rn[i] * rn[j] * rn[l] / d5;
        T3(o.DD2_Jr, i, j, l) = -6.0 * (KD(i, j) * rn[l] + KD(i, l) * rn[j] + KD(j, l) * rn[i]) / d5 +
                                30.0 * rn[i] * rn[j] * rn[l] / d7;
        for (int m = 0; m < 3; ++m) {
          T4(o.DD_Jr, i, j, m, l) = (-KD(i, j) * KD(l, m) + KD(i, l) * KD(j, m) + KD(j, l) * KD(i, m)) / d3 +
                                    15.0 * rn[i] * rn[j] * rn[l] * rn[m] / d7 -
                                    3.0 * (-KD(i, j) * rn[l] * rn[m] + KD(i, l) * rn[j] * rn[m] + KD(j, l) * rn[i] * rn[m] +
                                           KD(i, m) * rn[j] * rn[l] + KD(j, m) * rn[i] * rn[l] + KD(l, m) * rn[i] * rn[j]) / d5;
          /* forall(i,j,l,k): DDD2_Jg(i,j,k,l) with loop names (l->m here as k) */
          int k = m;
          T4(o.DDD2_Jr, i, j, k, l) = -6.0 * (KD(i, j) * KD(k, l) + KD(i, l) * KD(j, k) + KD(j, l) * KD(i, k)) / d5 -
                                      210.0 * rn[i] * rn[j] * rn[l] * rn[k] / d9 +
                                      30.0 * (-KD(i, j) * rn[l] * rn[k] + KD(i, l) * rn[j] * rn[k] + KD(j, l) * rn[i] * rn[k] +
                                              KD(i, k) * rn[j] * rn[l] + KD(j, k) * rn[i] * rn[l] + KD(l, k) * rn[i] * rn[j]) / d7;
        }
      }
    }
#undef KD
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double e = T3(g_EPS, j, k, l);
          T2(o.Ror, i, j) -= 0.5 * e * T3(o.D_Jr, i, l, k);
          T2(o.D2_Ror, i, j) -= 0.5 * e * T3(o.DD2_Jr, i, l, k);
          for (int q = 0; q < 3; ++q) T3(o.D_Ror, i, j, q) -= 0.5 * e * T4(o.DD_Jr, i, l, q, k);
        }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        T3(o.Kr, i, j, k) = 0.5 * (T3(o.D_Jr, i, j, k) + T3(o.D_Jr, i, k, j));
        T3(o.D2_Kr, i, j, k) = T3(o.DD2_Jr, i, j, k);
        for (int l = 0; l < 3; ++l) {
          T4(o.D_Kr, i, j, k, l) = 0.5 * (T4(o.DD_Jr, i, j, l, k) + T4(o.DD_Jr, i, k, l, j));
          T4(o.DD2_Kr, i, j, k, l) = T4(o.DDD2_Jr, i, j, k, l);
        }
      }
  pairacc_t A;
  memset(&A, 0, sizeof A);
  /* stokesian_Green.f90:64-100: TTPAB has no phix factor */
  accumulate_real(&A, &o, aI, aJ, 1.0, 1);
  finish_pair(&A, aI, TTP, RRP, TRP, QQP, GPQ, HPQ);
}

/* GRPY_YZD.f90:813-836 overlap_Regular_CORRECTION_IJ = regularised - free */
void orc_overlap_regular_correction(const double R[3], double aI, double aJ, double TT[9], double RR[9], double TR[9],
                                    double QQ[25], double GQ[15], double HQ[15]) {
  double a[9], b[9], c[9], d[25], e[15], f[15];
  orc_rp_regular(R, aI, aJ, TT, RR, TR, QQ, GQ, HQ);
  orc_rp_free_fts(R, aI, aJ, a, b, c, d, e, f);
  for (int q = 0; q < 9; ++q) {
    TT[q] -= a[q];
    RR[q] -= b[q];
    TR[q] -= c[q];
  }
  for (int q = 0; q < 25; ++q) QQ[q] -= d[q];
  for (int q = 0; q < 15; ++q) {
    GQ[q] -= e[q];
    HQ[q] -= f[q];
  }
}

/* ------------------------------------------------ Yamakawa (correction_method = 1)
 * GRPY_YZD.f90:326-786, kept literal: each block is built as the reference's
 * Cartesian tensor (CALC_RR, CALC_EPSR, CALC_UR, CALC_RRR, CALC_EPSRR,
 * CALC_t4D0/1/2 at tensors.f90:182-440) and projected with QQ_TR / GH_TR.
 * a1, a2 are the routine's (aaI, aaJ); "inside" means DIST <= |aaI - aaJ|. */
static void yam_tt_rr(const double R[3], double a1, double a2, int rr, double A[9]) {
  /* YAMAKAWA_TT_IJ :326-370, YAMAKAWA_RR_IJ :373-438 */
  const double d = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double w[3] = {R[0] / d, R[1] / d, R[2] / d};
  const double amax = a1 <= a2 ? a2 : a1;
  if (d <= fabs(a1 - a2)) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T2(A, i, j) = (i == j) ? (rr ? 1.0 / 8.0 / pow(amax, 3.0) : 1.0 / 6.0 / amax) : 0.0;
    return;
  }
  const double d2 = d * d, d3 = d2 * d, dm = a1 - a2;
  double pre, pu, prr;
  if (!rr) {
    pre = 1.0 / (6.0 * 32.0 * a1 * a2);
    pu = 16.0 * (a1 + a2) - pow(dm * dm + 3.0 * d2, 2.0) / d3;
    prr = 3.0 * pow(dm * dm - d2, 2.0) / d3;
  } else {
    const double i2 = a1 * a1, j2 = a2 * a2, i3 = i2 * a1, j3 = j2 * a2;
    pre = 1.0 / (8.0 * 64.0 * i3 * j3);
    pu = 5.0 * d3 - 27.0 * d * (i2 + j2) + 32.0 * (i3 + j3) - 9.0 * pow(i2 - j2, 2.0) / d -
         pow(dm, 4.0) * (i2 + 4.0 * a1 * a2 + j2) / d3;
    prr = 3.0 * pow(dm * dm - d2, 2.0) * ((i2 + 4.0 * a1 * a2 + j2) - d2) / d3;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T2(A, i, j) = pre * (pu * (i == j ? 1.0 : 0.0) + prr * w[i] * w[j]);
}

static void epsr(const double w[3], double E[9]) {  /* CALC_EPSR tensors.f90:201-218 */
  memset(E, 0, 9 * sizeof(double));
  T2(E, 0, 1) = w[2];
  T2(E, 1, 2) = w[0];
  T2(E, 2, 0) = w[1];
  T2(E, 1, 0) = -w[2];
  T2(E, 2, 1) = -w[0];
  T2(E, 0, 2) = -w[1];
}

static void yam_rt(const double R[3], double a1, double a2, double A[9]) {  /* YAMAKAWA_RT_IJ :467-505 */
  const double d = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double w[3] = {R[0] / d, R[1] / d, R[2] / d}, E[9];
  epsr(w, E);
  double c;
  if (d <= fabs(a1 - a2))
    c = (a1 <= a2) ? 0.0 : d / (8.0 * pow(a1, 3.0));
  else
    c = 1.0 / (16.0 * 8.0 * pow(a1, 3.0) * a2) * pow((a1 - a2) + d, 2.0) *
        (a2 * a2 + 2.0 * a2 * (a1 + d) - 3.0 * pow(a1 - d, 2.0)) / (d * d);
  for (int q = 0; q < 9; ++q) A[q] = c * E[q];
}

static void yam_td(const double R[3], double a1, double a2, double C[27]) {  /* YAMAKAWA_TD_IJ :582-640 */
  const double d = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double w[3] = {R[0] / d, R[1] / d, R[2] / d};
  double pur = 0.0, prrr = 0.0;
  if (d <= fabs(a1 - a2)) {
    if (a1 <= a2) pur = -3.0 * d / (20.0 * pow(a2, 3.0));
  } else {
    const double pre = 1.0 / (a1 * pow(a2, 3.0)), d2 = d * d, d4 = d2 * d2;
    pur = pre * (10.0 * d2 - 24.0 * a1 * d + 15.0 * (a1 * a1 - a2 * a2) - pow(a1 - a2, 5.0) * (a1 + 5.0 * a2) / d4) /
          320.0;
    prrr = pre * pow(pow(a1 - a2, 2.0) - d2, 2.0) * ((a1 - a2) * (a1 + 5.0 * a2) - d2) / (128.0 * d4);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        T3(C, i, j, k) = pur * (i == j ? w[k] : 0.0) + prrr * w[i] * w[j] * w[k];  /* t3UR, t3RRR */
}

static void yam_rd(const double R[3], double a1, double a2, double C[27]) {  /* YAMAKAWA_RD_IJ :512-556 */
  const double d = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double w[3] = {R[0] / d, R[1] / d, R[2] / d}, E[9];
  epsr(w, E);
  double c = 0.0;
  if (d > fabs(a1 - a2)) {
    const double d2 = d * d;
    c = -3.0 / (256.0 * pow(a1, 3.0) * pow(a2, 3.0) * d2 * d) * pow(pow(a1 - a2, 2.0) - d2, 2.0) *
        ((a1 * a1 + 4.0 * a1 * a2 + a2 * a2) - d2);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) T3(C, i, j, k) = c * T2(E, i, j) * w[k];  /* t3EPSRR */
}

static void yam_dd(const double R[3], double a1, double a2, double B[81]) {  /* YAMAKAWA_DD_IJ :665-765 */
  const double d = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double w[3] = {R[0] / d, R[1] / d, R[2] / d};
  double D0[81], D1[81], D2[81];
#define KD(i, j) ((i) == (j) ? 1.0 : 0.0)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          const double wwww = w[i] * w[j] * w[k] * w[l];
          T4(D0, i, j, k, l) = 1.5 * (w[i] * w[j] - KD(i, j) / 3.0) * (w[k] * w[l] - KD(k, l) / 3.0);
          T4(D1, i, j, k, l) = 0.5 * (KD(j, l) * w[i] * w[k] + KD(i, k) * w[j] * w[l] + KD(j, k) * w[i] * w[l] +
                                      KD(i, l) * w[j] * w[k] - 4.0 * wwww);
          T4(D2, i, j, k, l) = 0.5 * (KD(i, k) * KD(j, l) + KD(j, k) * KD(i, l) - KD(i, j) * KD(k, l) +
                                      KD(k, l) * w[i] * w[j] + KD(i, j) * w[k] * w[l] - 3.0 * wwww);
          T4(D2, i, j, k, l) -= T4(D1, i, j, k, l);
        }
#undef KD
  double c0, c1, c2;
  if (d <= fabs(a1 - a2)) {
    const double am = a1 <= a2 ? a2 : a1;
    c0 = c1 = c2 = 3.0 / (pow(am, 3.0) * 20.0);
  } else {
    const double i2 = a1 * a1, j2 = a2 * a2, i3 = i2 * a1, j3 = j2 * a2, d3 = d * d * d, d5 = d3 * d * d;
    const double m6 = pow(a1 - a2, 6.0) * (i2 + 6.0 * a1 * a2 + j2), m4 = pow(a1 - a2, 4.0) * (i2 + 4.0 * a1 * a2 + j2);
    c0 = 3.0 / (1280.0 * i3 * j3) * (3.0 * m6 / d5 - 10.0 * m4 / d3 + 32.0 * (i3 + j3) - 30.0 * (i2 + j2) * d + 5.0 * d3);
    c1 = 3.0 / (1280.0 * i3 * j3) *
         (-2.0 * m6 / d5 + 5.0 * m4 / d3 - 15.0 * pow(i2 - j2, 2.0) / d + 32.0 * (i3 + j3) - 25.0 * (i2 + j2) * d +
          5.0 * d3);
    c2 = 3.0 / (2560.0 * i3 * j3) *
         (m6 / d5 - 30.0 * pow(i2 - j2, 2.0) / d + 64.0 * (i3 + j3) - 40.0 * (i2 + j2) * d + 5.0 * d3);
  }
  for (int q = 0; q < 81; ++q) B[q] = c0 * D0[q] + c1 * D1[q] + c2 * D2[q];
}

/* GRPY_YZD.f90:768-786 YAMAKAWA_IJ: TR from RT with the pair reversed, transposed */
void orc_rp_yamakawa(const double R[3], double aI, double aJ, double TT[9], double RR[9], double TR[9], double QQ[25],
                     double GQ[15], double HQ[15]) {
  ensure_init();
  double mR[3] = {-R[0], -R[1], -R[2]}, RT0[9], G[27], H[27], Q[81];
  yam_tt_rr(R, aI, aJ, 0, TT);
  yam_tt_rr(R, aI, aJ, 1, RR);
  yam_rt(mR, aJ, aI, RT0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T2(TR, i, j) = T2(RT0, j, i);
  yam_td(R, aI, aJ, G);
  yam_rd(R, aI, aJ, H);
  yam_dd(R, aI, aJ, Q);
  gh_tr(GQ, G);
  gh_tr(HQ, H);
  qq_tr(QQ, Q);
}

/* GRPY_YZD.f90:789-811 overlap_Yamakawa_CORRECTION_IJ = Yamakawa - free */
void orc_overlap_yamakawa_correction(const double R[3], double aI, double aJ, double TT[9], double RR[9], double TR[9],
                                     double QQ[25], double GQ[15], double HQ[15]) {
  double a[9], b[9], c[9], d[25], e[15], f[15];
  orc_rp_yamakawa(R, aI, aJ, TT, RR, TR, QQ, GQ, HQ);
  orc_rp_free_fts(R, aI, aJ, a, b, c, d, e, f);
  for (int q = 0; q < 9; ++q) {
    TT[q] -= a[q];
    RR[q] -= b[q];
    TR[q] -= c[q];
  }
  for (int q = 0; q < 25; ++q) QQ[q] -= d[q];
  for (int q = 0; q < 15; ++q) {
    GQ[q] -= e[q];
    HQ[q] -= f[q];
  }
}

/* ------------------------------------------------ grand mobility */
/* per_GRPY_YZD.f90:13-200 PER_GRPERY_INV_FRI.  Computes the lattice
 * constants itself (CALC_LATTICE_INVERSE) and reports them in *lat.
 * correction_method: 0 none, 1 Yamakawa, 2 regularised (per_GRPY_YZD.f90:98-116). */
/* Explicit Ewald image sets (no counterpart in the reference, whose NR and NI
 * follow from EWS_ERR, csd_post.f90:158-183): nr >= 0 replaces NR and NI after
 * CALC_LATTICE_INVERSE; kmax2 > 0 keeps only the reciprocal vectors n with
 * |n|^2 <= kmax2 (a spherical cutoff inside the cube). */
static int g_nr_override = -1, g_ni_override = -1;
void orc_set_images(int nr, int ni, int kmax2) {
  g_nr_override = nr;
  g_ni_override = ni;
  g_kmax2 = kmax2;
}

int orc_per_grpery_inv_fri(int N, const double *pos, const double *radii, orc_lattice *lat, double ews_err,
                           int correction_method, double mu_f, double *APP, double *APQ, double *AQQ, int *n_overlap) {
  ensure_init();
  orc_calc_lattice_inverse(lat, ews_err, N, radii);
  if (g_nr_override >= 0) {
    lat->NR = g_nr_override;
    lat->NI = g_ni_override;
  }
  const long n11 = 11L * N;
  double *G = (double *)calloc((size_t)(n11 * n11), sizeof(double));
  if (!G) return -2;
#define GM(r, c) G[(r) + n11 * (long)(c)]
  int nov = 0;
  for (int al = 0; al < N; ++al) {
    double TTS[9], RRS[9], QQS[25];
    orc_rp_self(lat, radii[al], TTS, RRS, QQS);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        GM(3 * al + i, 3 * al + j) = TTS[i + 3 * j];
        GM(3 * N + 3 * al + i, 3 * N + 3 * al + j) = RRS[i + 3 * j];
      }
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) GM(6 * N + 5 * al + i, 6 * N + 5 * al + j) = QQS[i + 5 * j];
  }
  for (int al = 0; al < N; ++al)
    for (int be = al + 1; be < N; ++be) {
      double aI = radii[al], aJ = radii[be];
      double R[3], dmin;
      for (int i = 0; i < 3; ++i) R[i] = pos[3 * al + i] - pos[3 * be + i];
      orc_per_skew_correction(lat->LR, lat->LI, R, &dmin);
      double TT[9], RR[9], TR[9], QQ[25], GQ[15], HQ[15];
      orc_rp_ij(lat, R, aI, aJ, TT, RR, TR, QQ, GQ, HQ);
      if (correction_method != 0 && dmin <= aI + aJ) {
        double c1[9], c2[9], c3[9], c4[25], c5[15], c6[15];
        if (correction_method == 1)
          orc_overlap_yamakawa_correction(R, aI, aJ, c1, c2, c3, c4, c5, c6);
        else
          orc_overlap_regular_correction(R, aI, aJ, c1, c2, c3, c4, c5, c6);
        for (int q = 0; q < 9; ++q) {
          TT[q] += c1[q];
          RR[q] += c2[q];
          TR[q] += c3[q];
        }
        for (int q = 0; q < 25; ++q) QQ[q] += c4[q];
        for (int q = 0; q < 15; ++q) {
          GQ[q] += c5[q];
          HQ[q] += c6[q];
        }
        ++nov;
      }
      /* per_GRPY_YZD.f90:122-151 fills, then :160-188 scatter, both orientations */
      for (int pass = 0; pass < 2; ++pass) {
        int a = pass ? be : al, b = pass ? al : be;
        double sb = pass ? -1.0 : 1.0, sg = pass ? -1.0 : 1.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            GM(3 * a + i, 3 * b + j) = TT[i + 3 * j];
            GM(3 * a + i, 3 * N + 3 * b + j) = sb * TR[i + 3 * j];
            GM(3 * N + 3 * a + i, 3 * N + 3 * b + j) = RR[i + 3 * j];
            GM(3 * N + 3 * b + j, 3 * a + i) = sb * TR[i + 3 * j];
          }
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 5; ++j) {
            GM(3 * a + i, 6 * N + 5 * b + j) = sg * GQ[i + 3 * j];
            GM(3 * N + 3 * a + i, 6 * N + 5 * b + j) = HQ[i + 3 * j];
            GM(6 * N + 5 * b + j, 3 * a + i) = sg * GQ[i + 3 * j];
            GM(6 * N + 5 * b + j, 3 * N + 3 * a + i) = HQ[i + 3 * j];
          }
        for (int i = 0; i < 5; ++i)
          for (int j = 0; j < 5; ++j) GM(6 * N + 5 * a + i, 6 * N + 5 * b + j) = QQ[i + 5 * j];
      }
    }
  double sc = 1.0 / (PAI * mu_f);
  const long n6 = 6L * N, n5 = 5L * N;
  for (long c = 0; c < n6; ++c)
    for (long r = 0; r < n6; ++r) APP[r + n6 * c] = GM(r, c) * sc;
  for (long c = 0; c < n5; ++c)
    for (long r = 0; r < n6; ++r) APQ[r + n6 * c] = GM(r, n6 + c) * sc;
  for (long c = 0; c < n5; ++c)
    for (long r = 0; r < n5; ++r) AQQ[r + n5 * c] = GM(n6 + r, n6 + c) * sc;
#undef GM
  free(G);
  if (n_overlap) *n_overlap = nov;
  return 0;
}

/* ------------------------------------------------ lubrication */
/* hydro_lub.f90:400-525 tables (SAVEd after the first read) */
typedef struct {
  int nabc, ngh, nm;
  double abc[39][11];
  double gh[47][7];
  double m[47][4];
} orc_lub_tables;

static int read_table(const char *dir, const char *name, int rows, int cols, double *out) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[1024];
  if (!fgets(line, sizeof line, f)) {
    fclose(f);
    return -1;
  }
  for (int i = 0; i < rows * cols; ++i)
    if (fscanf(f, "%lf", &out[i]) != 1) {
      fclose(f);
      return -1;
    }
  fclose(f);
  return 0;
}
/* hydro_lub.f90:529-596 readlubabc / readlubgh / readlubm */
int orc_lub_tables_read(const char *dir, orc_lub_tables *t) {
  t->nabc = 39;
  t->ngh = 47;
  t->nm = 47;
  if (read_table(dir, "r2babc.dat", 39, 11, &t->abc[0][0])) return -1;
  if (read_table(dir, "r2bgh.dat", 47, 7, &t->gh[0][0])) return -1;
  if (read_table(dir, "r2bm.dat", 47, 4, &t->m[0][0])) return -1;
  return 0;
}
int orc_lub_tables_size(void) { return (int)sizeof(orc_lub_tables); }

/* hydro_lub.f90:310-398 lubcalc / :400-525 lubtabcalc.  out[19] =
 * x11a,x12a,y11a,y12a,y11b,y12b,x11c,x12c,y11c,y12c,x11g,x12g,y11g,y12g,y11h,y12h,xm,ym,zm */
void orc_lubcalc(double r, const orc_lub_tables *t, double o[19]) {
  if (r < 2.1) {
    double xi = r - 2.0, invxi = 1.0 / xi, li = log(invxi), xl = xi * li;
    double x11a = -1.23041 + 0.25 * invxi + 1.8918 * xi + 9.0 * li / 40.0 + 3.0 * xl / 112.0;
    double x12a = -x11a + 0.00312 - 0.0011 * xi;
    double y11a = -0.39394 + 0.95665 * xi + li / 6.0;
    double y12a = -y11a + 0.004636 - 0.007049 * xi;
    double y11b = 0.408286 - 0.84055 * xi - li / 6.0 - xl / 12.0;
    double x11c = 0.0479 + 0.12494 * xi - xl / 6.0;
    double x12c = -x11c + 0.016869 + 0.049536 * xi;
    double y11c = -0.605434 + 0.939139 * xi + 4.0 * li / 15.0 + 94.0 * xl / 375.0;
    double y12c = li / 15.0 + 62.0 * xl / 375.0 - 0.219032 + 0.332843 * xi;
    double y12b = -0.405978 + 0.833042 * xi + li / 6.0 + xl / 12.0;
    double x11g = -1.16897 + 1.47882 * xi + invxi / 4.0 + 9.0 * li / 40.0 + 39.0 * xl / 280.0;
    double x12g = -x11g + 0.01 - 0.00167 * xi;
    double y11g = -0.2041 + 0.44226 * xi + li / 12.0 + xl / 24.0;
    double y12g = -y11g + 0.012265 - 0.02757 * xi;
    double y11h = -0.143777 + 0.264207 * xi + li / 30.0 + 137.0 * xl / 1500.0;
    double y12h = -0.298166 + 0.534123 * xi + 2.0 * li / 15.0 + 113.0 * xl / 1500.0;
    double xm = invxi / 6.0 + 3.0 * li / 20.0 + 47.0 * xl / 280.0 - 0.740815 + 0.706802 * xi;
    double ym = -0.193518 + 0.553883 * xi + li / 12.0 - 29.0 * xl / 500.0;
    double zm = 0.00645755 - 0.021142 * xi;
    double v[19] = {x11a, x12a, y11a, y12a, y11b, y12b, x11c, x12c, y11c, y12c,
                    x11g, x12g, y11g, y12g, y11h, y12h, xm, ym, zm};
    memcpy(o, v, sizeof v);
    return;
  }
  for (int q = 0; q < 19; ++q) o[q] = 0.0;
  for (int i = 0; i < t->nabc - 1; ++i) {
    double t0 = t->abc[i][0], t1 = t->abc[i + 1][0];
    if (t0 <= r && t1 > r) {
      for (int c = 0; c < 10; ++c)
        o[c] = t->abc[i][c + 1] + (r - t0) * (t->abc[i + 1][c + 1] - t->abc[i][c + 1]) / (t1 - t0);
      break;
    }
  }
  for (int i = 0; i < t->ngh - 1; ++i) {
    double t0 = t->gh[i][0], t1 = t->gh[i + 1][0];
    if (t0 <= r && t1 > r) {
      for (int c = 0; c < 6; ++c)
        o[10 + c] = t->gh[i][c + 1] + (r - t0) * (t->gh[i + 1][c + 1] - t->gh[i][c + 1]) / (t1 - t0);
      break;
    }
  }
  for (int i = 0; i < t->nm - 1; ++i) {
    double t0 = t->m[i][0], t1 = t->m[i + 1][0];
    if (t0 <= r && t1 > r) {
      for (int c = 0; c < 3; ++c)
        o[16 + c] = (t->m[i][c + 1] + (r - t0) * (t->m[i + 1][c + 1] - t->m[i][c + 1]) / (t1 - t0)) / 2.0;
      break;
    }
  }
}

/* tensors.f90:936-950 DD_transY2 */
static void dd_trans(double B[25], const double Bc[81]) { qq_tr(B, Bc); }

/* hydro_lub.f90:22-308 lubmxcalc.  rfu(6N,6N), rfe(6N,5N), rse(5N,5N).
 * Returns the number of contact reflections (s<=2). */
int orc_lubmxcalc(int N, int Nb, int is_period, const double *pos, const double *radii, const double LR[9],
                  const double LI[9], double mu_f, const orc_lub_tables *tab, double *rfu, double *rfe, double *rse) {
  ensure_init();
  const long n6 = 6L * N, n5 = 5L * N;
  double *APPl = (double *)calloc((size_t)(n6 * n6), sizeof(double));
  double *AQPl = (double *)calloc((size_t)(n5 * n6), sizeof(double));
  double *AQQl = (double *)calloc((size_t)(n5 * n5), sizeof(double));
  if (!APPl || !AQPl || !AQQl) return -2;
#define P(r, c) APPl[(r) + n6 * (long)(c)]
#define QP(r, c) AQPl[(r) + n5 * (long)(c)]
#define QQ(r, c) AQQl[(r) + n5 * (long)(c)]
  int ncontact = 0;
  for (int al = 0; al < N; ++al)
    for (int be = al + 1; be < N; ++be) {
      if (al > N - Nb - 1) continue;
      double rx[3], r;
      for (int i = 0; i < 3; ++i) rx[i] = pos[3 * al + i] - pos[3 * be + i];
      r = sqrt(rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2]);
      if (is_period) orc_per_skew_correction(LR, LI, rx, &r);
      double r_ave = 0.5 * (radii[al] + radii[be]);
      double s = r / r_ave;
      double v[19];
      if (s <= 2.0) {
        ++ncontact;
        s = 4.00001 - s;
        orc_lubcalc(s, tab, v);
      } else if (s > 4.0) {
        continue;
      } else {
        orc_lubcalc(s, tab, v);
      }
      double x11a = v[0], x12a = v[1], y11a = v[2], y12a = v[3], y11b = v[4], y12b = v[5], x11c = v[6], x12c = v[7],
             y11c = v[8], y12c = v[9], x11g = v[10], x12g = v[11], y11g = v[12], y12g = v[13], y11h = v[14],
             y12h = v[15], xm = v[16], ym = v[17], zm = v[18];
      double e[3] = {rx[0] / r, rx[1] / r, rx[2] / r};
      double a11[9], a12[9], b11[9], b12[9], c11[9], c12[9];
      double g11[27], g12[27], h11[27], h12[27], ml[81];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double kd = T2(g_KD, i, j), ee = e[i] * e[j];
          double pe = T3(g_EPS, i, j, 0) * e[0] + T3(g_EPS, i, j, 1) * e[1] + T3(g_EPS, i, j, 2) * e[2];
          T2(a11, i, j) = x11a * ee + y11a * (kd - ee);
          T2(a12, i, j) = x12a * ee + y12a * (kd - ee);
          T2(b11, i, j) = y11b * pe;
          T2(b12, i, j) = y12b * pe;
          T2(c11, i, j) = x11c * ee + y11c * (kd - ee);
          T2(c12, i, j) = x12c * ee + y12c * (kd - ee);
        }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            double kij = T2(g_KD, i, j), kjk = T2(g_KD, j, k), kik = T2(g_KD, i, k);
            double base_x = (e[i] * e[j] - kij / 3.0) * e[k];
            double base_y = e[i] * kjk + e[j] * kik - 2 * e[i] * e[j] * e[k];
            T3(g11, i, j, k) = x11g * base_x + y11g * base_y;
            T3(g12, i, j, k) = x12g * base_x + y12g * base_y;
            double pik = T3(g_EPS, i, k, 0) * e[0] + T3(g_EPS, i, k, 1) * e[1] + T3(g_EPS, i, k, 2) * e[2];
            double pjk = T3(g_EPS, j, k, 0) * e[0] + T3(g_EPS, j, k, 1) * e[1] + T3(g_EPS, j, k, 2) * e[2];
            T3(h11, i, j, k) = y11h * (pik * e[j] + pjk * e[i]);
            T3(h12, i, j, k) = y12h * (pik * e[j] + pjk * e[i]);
          }
#define KD(a, b) T2(g_KD, a, b)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l)
              T4(ml, i, j, k, l) =
                  xm * (3.0 * (e[i] * e[j] - KD(i, j) / 3.0) * (e[k] * e[l] - KD(k, l) / 3.0) / 2.0) +
                  ym * (e[i] * KD(j, l) * e[k] + e[j] * KD(i, l) * e[k] + e[i] * KD(j, k) * e[l] + e[j] * KD(i, k) * e[l] -
                        4.0 * e[i] * e[j] * e[k] * e[l]) / 2.0 +
                  zm * (KD(i, k) * KD(j, l) + KD(j, k) * KD(i, l) - KD(i, j) * KD(k, l) + e[i] * e[j] * KD(k, l) +
                        KD(i, j) * e[k] * e[l] - e[i] * KD(j, l) * e[k] - e[j] * KD(i, l) * e[k] -
                        e[i] * KD(j, k) * e[l] - e[j] * KD(i, k) * e[l] + e[i] * e[j] * e[k] * e[l]) / 2.0;
#undef KD
      double mmx[25];
      dd_trans(mmx, ml);
      /* (2,2) and (2,1) blocks by the sign rules of hydro_lub.f90:144-159 */
      double g11y[15], g12y[15], h11y[15], h12y[15];
      rtd_front(g11y, g11);
      rtd_front(g12y, g12);
      rtd_front(h11y, h11);
      rtd_front(h12y, h12);
      const long TA = 3L * al, TB = 3L * be, RA = 3L * N + 3L * al, RB = 3L * N + 3L * be;
      const long QA = 5L * al, QB = 5L * be;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          P(TA + i, TA + j) += T2(a11, i, j);
          P(TB + i, TB + j) += T2(a11, i, j);
          P(RA + i, RA + j) += T2(c11, i, j);
          P(RB + i, RB + j) += T2(c11, i, j);
          P(RA + i, TA + j) += T2(b11, i, j);
          P(RB + i, TB + j) += -T2(b11, i, j);
          P(TA + i, TB + j) += T2(a12, i, j);
          P(RA + i, RB + j) += T2(c12, i, j);
          P(RA + i, TB + j) += T2(b12, i, j);
          P(RB + i, TA + j) += -T2(b12, i, j);
        }
      for (int I = 0; I < 5; ++I)
        for (int c = 0; c < 3; ++c) {
          QP(QA + I, TA + c) += g11y[I + 5 * c];
          QP(QB + I, TB + c) += -g11y[I + 5 * c];
          QP(QA + I, 3L * N + TA + c) += h11y[I + 5 * c];
          QP(QB + I, 3L * N + TB + c) += h11y[I + 5 * c];
          QP(QA + I, TB + c) += g12y[I + 5 * c];
          QP(QB + I, TA + c) += -g12y[I + 5 * c];
          QP(QA + I, 3L * N + TB + c) += h12y[I + 5 * c];
          QP(QB + I, 3L * N + TA + c) += h12y[I + 5 * c];
        }
      for (int I = 0; I < 5; ++I)
        for (int J = 0; J < 5; ++J) {
          QQ(QA + I, QA + J) += mmx[I + 5 * J];
          QQ(QB + I, QB + J) += mmx[I + 5 * J];
          QQ(QA + I, QB + J) += mmx[I + 5 * J];
          QQ(QB + I, QA + J) += mmx[I + 5 * J];
        }
    }
  /* hydro_lub.f90:296-302 */
  for (long r = 0; r < 3L * N; ++r)
    for (long c = 0; c < 3L * N; ++c) P(r, 3L * N + c) = P(3L * N + c, r);
  for (long I = 0; I < n6 - 1; ++I)
    for (long J = I + 1; J < n6; ++J) P(J, I) = P(I, J);
  double sc = 6.0 * (PAI * mu_f);
  for (long q = 0; q < n6 * n6; ++q) rfu[q] = APPl[q] * sc;
  for (long c = 0; c < n5; ++c)
    for (long r = 0; r < n6; ++r) rfe[r + n6 * c] = QP(c, r) * sc;
  for (long q = 0; q < n5 * n5; ++q) rse[q] = AQQl[q] * sc;
#undef P
#undef QP
#undef QQ
  free(APPl);
  free(AQPl);
  free(AQQl);
  return ncontact;
}

/* ------------------------------------------------ stepping helpers */
/* csd_post.f90:497-522 calc_uo_bg with INIT_u_bg (:442-479): u = u_bg + (OmegaT+E) x,
 * omega = omega_bg.  Inputs are the background fields as the reference
 * holds them after INIT_u_bg. */
void orc_calc_uo_bg(int N, const double *conf, const double u_bg[3], const double omega_bg[3], const double Eij[9],
                    double *uo_bg) {
  ensure_init();
  double omT[9], U[9];
  for (int I = 0; I < 3; ++I)
    for (int J = 0; J < 3; ++J)
      omT[I + 3 * J] = T3(g_EPS, J, I, 0) * omega_bg[0] + T3(g_EPS, J, I, 1) * omega_bg[1] + T3(g_EPS, J, I, 2) * omega_bg[2];
  for (int q = 0; q < 9; ++q) U[q] = omT[q] + Eij[q];
  for (int i = 0; i < N; ++i) {
    for (int r = 0; r < 3; ++r)
      uo_bg[3 * i + r] = u_bg[r] + (U[r] * conf[3 * i] + U[r + 3] * conf[3 * i + 1] + U[r + 6] * conf[3 * i + 2]);
    for (int r = 0; r < 3; ++r) uo_bg[3 * N + 3 * i + r] = omega_bg[r];
  }
}

/* csd_post.f90:383-414 pos_collision_judge (raw distance, no minimum image) */
int orc_pos_collision(int N, int Nb, const double *conf, const double *radii, const int *floc) {
  for (int a = 0; a < N - Nb; ++a)
    for (int b = a + 1; b < N; ++b) {
      if (floc[a] == 0 && floc[b] == 0) continue;
      double d0 = conf[3 * a] - conf[3 * b], d1 = conf[3 * a + 1] - conf[3 * b + 1], d2 = conf[3 * a + 2] - conf[3 * b + 2];
      double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      if (r <= radii[a] + radii[b]) return 1;
    }
  return 0;
}

/* ------------------------------------------------ free space and wall */
/* GRPY_YZD.f90:14-264 GRPERY_INV_FRI (FTS, IsPeriod = false): self blocks
 * ROTNE_PRAGER_TT/RR_SELF and DD_SELF_Y2 (:275-315), pair blocks
 * ROTNE_PRAGER_IJ_FTS or, for r <= aI + aJ with correction_method 2,
 * Regular_ROTNE_PRAGER_IJ, scattered as :166-239, divided by pi mu_f (:255-257). */
int orc_grpery_inv_fri(int N, const double *pos, const double *radii, int correction_method, double mu_f,
                       double *APP, double *APQ, double *AQQ, int *n_overlap) {
  ensure_init();
  const long n11 = 11L * N;
  double *G = (double *)calloc((size_t)(n11 * n11), sizeof(double));
  if (!G) return -2;
#define GM(r, c) G[(r) + n11 * (long)(c)]
  int nov = 0;
  for (int al = 0; al < N; ++al) {
    double a = radii[al], B1c[81], B1[25];
    for (int i = 0; i < 3; ++i) {
      GM(3 * al + i, 3 * al + i) = 1.0 / (6.0 * a);
      GM(3 * N + 3 * al + i, 3 * N + 3 * al + i) = 1.0 / (8.0 * a * a * a);
    }
    for (int q = 0; q < 81; ++q) B1c[q] = 3.0 / (20.0 * a * a * a) * g_II[q];
    dd_trans(B1, B1c);
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) GM(6 * N + 5 * al + i, 6 * N + 5 * al + j) = B1[i + 5 * j];
  }
  for (int al = 0; al < N; ++al)
    for (int be = al + 1; be < N; ++be) {
      double aI = radii[al], aJ = radii[be], R[3];
      for (int i = 0; i < 3; ++i) R[i] = pos[3 * al + i] - pos[3 * be + i];
      double dmin = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      double TT[9], RR[9], TR[9], QQ[25], GQ[15], HQ[15];
      if (dmin <= aI + aJ && correction_method != 0) {
        if (correction_method == 1)
          orc_rp_yamakawa(R, aI, aJ, TT, RR, TR, QQ, GQ, HQ);
        else
          orc_rp_regular(R, aI, aJ, TT, RR, TR, QQ, GQ, HQ);
        ++nov;
      } else {
        orc_rp_free_fts(R, aI, aJ, TT, RR, TR, QQ, GQ, HQ);
      }
      for (int pass = 0; pass < 2; ++pass) {
        int a = pass ? be : al, b = pass ? al : be;
        double sg = pass ? -1.0 : 1.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            GM(3 * a + i, 3 * b + j) = TT[i + 3 * j];
            GM(3 * a + i, 3 * N + 3 * b + j) = sg * TR[i + 3 * j];
            GM(3 * N + 3 * a + i, 3 * N + 3 * b + j) = RR[i + 3 * j];
            GM(3 * N + 3 * b + j, 3 * a + i) = sg * TR[i + 3 * j];
          }
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 5; ++j) {
            GM(3 * a + i, 6 * N + 5 * b + j) = sg * GQ[i + 3 * j];
            GM(3 * N + 3 * a + i, 6 * N + 5 * b + j) = HQ[i + 3 * j];
            GM(6 * N + 5 * b + j, 3 * a + i) = sg * GQ[i + 3 * j];
            GM(6 * N + 5 * b + j, 3 * N + 3 * a + i) = HQ[i + 3 * j];
          }
        for (int i = 0; i < 5; ++i)
          for (int j = 0; j < 5; ++j) GM(6 * N + 5 * a + i, 6 * N + 5 * b + j) = QQ[i + 5 * j];
      }
    }
  double sc = 1.0 / (PAI * mu_f);
  const long n6 = 6L * N, n5 = 5L * N;
  for (long c = 0; c < n6; ++c)
    for (long r = 0; r < n6; ++r) APP[r + n6 * c] = GM(r, c) * sc;
  for (long c = 0; c < n5; ++c)
    for (long r = 0; r < n6; ++r) APQ[r + n6 * c] = GM(r, n6 + c) * sc;
  for (long c = 0; c < n5; ++c)
    for (long r = 0; r < n5; ++r) AQQ[r + n5 * c] = GM(n6 + r, n6 + c) * sc;
#undef GM
  free(G);
  if (n_overlap) *n_overlap = nov;
  return 0;
}

/* Wall_interaction.f90:29-149 Wall_GRPERY_INV_FRI: TT blocks of each particle
 * with the images (x, y, -z) of itself (z <= 10) and of its partners (either
 * z <= 10), ROTNE_PRAGER_TT_IJ with the third column negated
 * (Wall_P_Right_tensor, :176-185), block (beta, alpha) = block (alpha, beta). */
void orc_wall_grpery_inv_fri(int N, const double *pos, const double *radii, double mu_f, double *TPP) {
  ensure_init();
  const long n3 = 3L * N;
  for (long q = 0; q < n3 * n3; ++q) TPP[q] = 0.0;
  double TT[9], RR[9], TR[9], QQ[25], GQ[15], HQ[15];
  for (int al = 0; al < N; ++al) {
    if (!(pos[3 * al + 2] <= 10.0)) continue;
    double R[3] = {0.0, 0.0, 2.0 * pos[3 * al + 2]};
    orc_rp_free_fts(R, radii[al], radii[al], TT, RR, TR, QQ, GQ, HQ);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) TPP[(3 * al + i) + n3 * (3 * al + j)] = (j == 2 ? -1.0 : 1.0) * TT[i + 3 * j];
  }
  for (int al = 0; al < N; ++al)
    for (int be = al + 1; be < N; ++be) {
      if (!(pos[3 * al + 2] <= 10.0 || pos[3 * be + 2] <= 10.0)) continue;
      double R[3] = {pos[3 * al] - pos[3 * be], pos[3 * al + 1] - pos[3 * be + 1], pos[3 * al + 2] + pos[3 * be + 2]};
      orc_rp_free_fts(R, radii[al], radii[be], TT, RR, TR, QQ, GQ, HQ);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = (j == 2 ? -1.0 : 1.0) * TT[i + 3 * j];
          TPP[(3 * al + i) + n3 * (3 * be + j)] = v;
          TPP[(3 * be + i) + n3 * (3 * al + j)] = v;
        }
    }
  double sc = 1.0 / (PAI * mu_f);
  for (long q = 0; q < n3 * n3; ++q) TPP[q] *= sc;
}
