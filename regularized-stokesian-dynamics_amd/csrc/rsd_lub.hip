// rsd_lub.hip -- pairwise lubrication resistance (lubmxcalc, hydro_lub.f90:22-308).
//
// One 64-lane workgroup per particle; see lub_kernel.
#include <hip/hip_runtime.h>

#include "rsd_lub.h"
#include "rsd_lubmath.h"
#include "rsd_prof.h"

namespace rsd {

using namespace rsdm;

__device__ __forceinline__ double lub_skew(const LubParams& p, double& x, double& y, double& z) {
  double X[3];
  for (int i = 0; i < 3; ++i) {
    X[i] = p.LI[i] * x + p.LI[i + 3] * y + p.LI[i + 6] * z;
    X[i] = fmod(X[i], 1.0);
    if (X[i] < 0.0) X[i] += 1.0;
  }
  double RC[3];
  for (int i = 0; i < 3; ++i) RC[i] = p.LR[i] * X[0] + p.LR[i + 3] * X[1] + p.LR[i + 6] * X[2];
  double D = RC[0] * RC[0] + RC[1] * RC[1] + RC[2] * RC[2];
  double bx = RC[0], by = RC[1], bz = RC[2];
  for (int m = 1; m < 8; ++m) {
    const int m1 = (m >> 2) & 1, m2 = (m >> 1) & 1, m3 = m & 1;
    double d0 = RC[0] - (p.LR[0] * m1 + p.LR[3] * m2 + p.LR[6] * m3);
    double d1 = RC[1] - (p.LR[1] * m1 + p.LR[4] * m2 + p.LR[7] * m3);
    double d2 = RC[2] - (p.LR[2] * m1 + p.LR[5] * m2 + p.LR[8] * m3);
    double d = d0 * d0 + d1 * d1 + d2 * d2;
    if (d < D) {
      D = d;
      bx = d0;
      by = d1;
      bz = d2;
    }
  }
  x = bx;
  y = by;
  z = bz;
  return sqrt(D);
}

constexpr int kSelfAcc = 6 + 6 + 3 + 15 + 15 + 15;  // a, c, b, g, h, m
constexpr int kCrossVals = 6 + 6 + 3 + 15 + 15 + 15;  // a, c, b, g, h, m of one (alpha, beta) pair
constexpr int kQueue = 128;
constexpr int kScatterPairs = 4;  // cross pairs whose read-modify-writes are in flight together

// Index of (i, j) in the symmetric 3x3 order 00,11,22,01,02,12.
__device__ __forceinline__ int sym3(int i, int j) {
  return i == j ? i : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5));
}
// Index of (I, J) in the packed upper 5x5 order (row by row from the diagonal).
__device__ __forceinline__ int packed5(int I, int J) {
  const int lo = I < J ? I : J, hi = I < J ? J : I;
  return 5 * lo - lo * (lo - 1) / 2 + (hi - lo);
}
// The off-diagonal (i, j) pairs of a 3x3 block.
__device__ __forceinline__ void offdiag3(int k, int& i, int& j) {
  i = k >> 1;
  const int t = k & 1;
  j = (i == 0) ? 1 + t : (i == 1 ? 2 * t : t);
}

// Entry e (0..169) of the cross blocks of the pair (al < be), values v =
// scaled a, c, b, g, h, m (hydro_lub.f90:182-302): the full symmetric TT and
// RR blocks at (al,be) and (be,al), the TR/RT couplings eps.b, the PQ
// blocks g and h, and the QQ block m.  Every entry is a distinct matrix
// element, so the 170 read-modify-writes of one pair can run in parallel.
__device__ __forceinline__ double* cross_entry(const LubParams& p, const double* v, int al, int be, int e,
                                               double& val) {
  const int N = p.N;
  double* M;
  int ld, r, c;
  if (e < 36) {  // a (TT) and c (RR), both orientations
    const int t = e / 18, q = e % 18, o = q / 9, i = (q % 9) / 3, j = q % 3;
    const int off = t ? 3 * N : 0;
    M = p.PP;
    ld = p.ldPP;
    r = off + 3 * (o ? be : al) + i;
    c = off + 3 * (o ? al : be) + j;
    val = v[6 * t + sym3(i, j)];
  } else if (e < 60) {  // B12(i,j) = eps_ijk b_k: RT(al,be)=B, RT(be,al)=-B, TR = RT^T
    const int k = e - 36, var = k / 6;
    int i, j;
    offdiag3(k % 6, i, j);
    const int kk = 3 - i - j;  // the third index
    const double sgn = ((j - i + 3) % 3 == 1) ? 1.0 : -1.0;  // eps_ijk for (i,j,kk)
    const double Bij = sgn * v[12 + kk];
    M = p.PP;
    ld = p.ldPP;
    if (var == 0) { r = 3 * N + 3 * al + i; c = 3 * be + j; val = Bij; }
    else if (var == 1) { r = 3 * be + j; c = 3 * N + 3 * al + i; val = Bij; }
    else if (var == 2) { r = 3 * N + 3 * be + i; c = 3 * al + j; val = -Bij; }
    else { r = 3 * al + j; c = 3 * N + 3 * be + i; val = -Bij; }
  } else if (e < 120) {  // PQ: g and h
    const int k = e - 60, var = k / 15, I = (k % 15) % 5, kk = (k % 15) / 5;
    M = p.PQ;
    ld = p.ldPQ;
    const double g = v[15 + I + 5 * kk], h = v[30 + I + 5 * kk];
    if (var == 0) { r = 3 * be + kk; c = 5 * al + I; val = g; }
    else if (var == 1) { r = 3 * al + kk; c = 5 * be + I; val = -g; }
    else if (var == 2) { r = 3 * N + 3 * be + kk; c = 5 * al + I; val = h; }
    else { r = 3 * N + 3 * al + kk; c = 5 * be + I; val = h; }
  } else {  // QQ: m, both orientations
    const int k = e - 120, o = k / 25, I = (k % 25) / 5, J = k % 5;
    M = p.QQ;
    ld = p.ldQQ;
    r = 5 * (o ? be : al) + I;
    c = 5 * (o ? al : be) + J;
    val = v[45 + packed5(I, J)];
  }
  return M + (size_t)r + (size_t)c * ld;
}
constexpr int kCrossEntries = 170;
constexpr int kCrossPPEntries = 60;  // cross entries below this are PP's, the rest PQ's and QQ's

// Entry e (0..84) of the self blocks of particle g from the reduced, scaled
// accumulators v (a, c, b, g, h, m as above).
__device__ __forceinline__ double* self_entry(const LubParams& p, const double* v, int g, int e, double& val) {
  const int N = p.N;
  double* M;
  int ld, r, c;
  if (e < 18) {
    const int t = e / 9, i = (e % 9) / 3, j = e % 3, off = t ? 3 * N : 0;
    M = p.PP;
    ld = p.ldPP;
    r = off + 3 * g + i;
    c = off + 3 * g + j;
    val = v[6 * t + sym3(i, j)];
  } else if (e < 30) {
    const int k = e - 18, var = k / 6;
    int i, j;
    offdiag3(k % 6, i, j);
    const int kk = 3 - i - j;
    const double sgn = ((j - i + 3) % 3 == 1) ? 1.0 : -1.0;
    const double Bij = sgn * v[12 + kk];
    M = p.PP;
    ld = p.ldPP;
    if (var == 0) { r = 3 * N + 3 * g + i; c = 3 * g + j; }
    else { r = 3 * g + j; c = 3 * N + 3 * g + i; }
    val = Bij;
  } else if (e < 60) {
    const int k = e - 30, var = k / 15, I = (k % 15) % 5, cc = (k % 15) / 5;
    M = p.PQ;
    ld = p.ldPQ;
    r = (var ? 3 * N : 0) + 3 * g + cc;
    c = 5 * g + I;
    val = v[(var ? 30 : 15) + I + 5 * cc];
  } else {
    const int k = e - 60, I = k / 5, J = k % 5;
    M = p.QQ;
    ld = p.ldQQ;
    r = 5 * g + I;
    c = 5 * g + J;
    val = v[45 + packed5(I, J)];
  }
  return M + (size_t)r + (size_t)c * ld;
}
constexpr int kSelfEntries = 85;
constexpr int kSelfPPEntries = 30;  // self entries below this are PP's, the rest PQ's and QQ's

// One 64-lane workgroup per particle gamma.  The lanes sweep the partners
// delta in order, 64 at a time, and queue the ones inside the lubrication
// range (s <= 4 after the contact reflection) in LDS, keeping the sweep
// order.  Each batch of 64 queued pairs is then evaluated one pair per lane
// (tables, projections, blocks), so the heavy code runs about once per
// particle instead of once per 64 partners.  Each pair is oriented as the
// reference orients it (lower index minus higher index, then
// PER_SKEW_CORRECTION), so the (1,1)/(2,2) self blocks and the (1,2)/(2,1)
// cross blocks get the reference's signs.  Self contributions are reduced
// across the wave in a fixed order; the cross blocks of (alpha < beta) are
// staged in LDS by alpha's workgroup and scattered by all its lanes, a few
// pairs per memory round trip.  No
// atomics on the matrices: the result is bitwise repeatable.
// MODE: kLubDense writes all three blocks; kLubProducts adds PP and forms PQ ein,
// QQ ein (LubParams::ein); kLubTranspose forms PQ^T xt only (LubParams::xt)
// kLubBsr writes the block-sparse rows (LubParams::bsr_*).
enum LubMode : int { kLubDense = 0, kLubProducts = 1, kLubTranspose = 2, kLubBsr = 3 };

// The 11x11 block (gamma, delta) of the grand lubrication resistance in
// gamma's and delta's own dof order (T 0-2, R 3-5, Q 6-10), column-major,
// from a pair's scaled a, c, b, g, h, m (v, the cross_entry layout).  sigma
// is +1 when gamma is the pair's lower index.  Read off cross_entry's
// placements (hydro_lub.f90:182-302 and rfe^T for the Q rows):
//   TT a, RR c, RT = TR = sigma eps.b, TQ = -sigma g, RQ = h,
//   QT = sigma g^T, QR = h^T, QQ m.
__device__ __forceinline__ void bsr_cross_block(const double* v, double sigma, double* __restrict__ blk) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      blk[i + 11 * j] = v[sym3(i, j)];
      blk[(3 + i) + 11 * (3 + j)] = v[6 + sym3(i, j)];
      const double Bij = (i == j) ? 0.0 : (((j - i + 3) % 3 == 1) ? 1.0 : -1.0) * v[12 + (3 - i - j)];
      blk[(3 + i) + 11 * j] = sigma * Bij;
      blk[i + 11 * (3 + j)] = sigma * Bij;
    }
  for (int k = 0; k < 3; ++k)
    for (int I = 0; I < 5; ++I) {
      const double g = v[15 + I + 5 * k], h = v[30 + I + 5 * k];
      blk[k + 11 * (6 + I)] = -sigma * g;
      blk[(3 + k) + 11 * (6 + I)] = h;
      blk[(6 + I) + 11 * k] = sigma * g;
      blk[(6 + I) + 11 * (3 + k)] = h;
    }
  for (int I = 0; I < 5; ++I)
    for (int J = 0; J < 5; ++J) blk[(6 + I) + 11 * (6 + J)] = v[45 + packed5(I, J)];
}

// The self block (gamma, gamma) from the reduced accumulators (b and g with
// their signs already): TT a, RR c, RT = eps.b, TR = RT^T, TQ = g, RQ = h,
// QT = g^T, QR = h^T, QQ m (self_entry's placements).
__device__ __forceinline__ void bsr_self_block(const double* v, double* __restrict__ blk) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      blk[i + 11 * j] = v[sym3(i, j)];
      blk[(3 + i) + 11 * (3 + j)] = v[6 + sym3(i, j)];
      const double Bij = (i == j) ? 0.0 : (((j - i + 3) % 3 == 1) ? 1.0 : -1.0) * v[12 + (3 - i - j)];
      blk[(3 + i) + 11 * j] = Bij;
      blk[j + 11 * (3 + i)] = Bij;
    }
  for (int k = 0; k < 3; ++k)
    for (int I = 0; I < 5; ++I) {
      const double g = v[15 + I + 5 * k], h = v[30 + I + 5 * k];
      blk[k + 11 * (6 + I)] = g;
      blk[(3 + k) + 11 * (6 + I)] = h;
      blk[(6 + I) + 11 * k] = g;
      blk[(6 + I) + 11 * (3 + k)] = h;
    }
  for (int I = 0; I < 5; ++I)
    for (int J = 0; J < 5; ++J) blk[(6 + I) + 11 * (6 + J)] = v[45 + packed5(I, J)];
}

template <int MODE>
__global__ __launch_bounds__(64) void lub_kernel(LubParams p, const double* __restrict__ pos,
                                                 const double* __restrict__ radii, const LubTables* __restrict__ tab,
                                                 int* __restrict__ contacts) {
  __shared__ int q_idx[kQueue];
  __shared__ double q_geo[kQueue][4];  // rx, ry, rz (lo - hi, minimum image), s (reflected)
  __shared__ double stage[64][kCrossVals + 1];
  __shared__ int stage_be[64];
  const int gm = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = p.N;
  const double sc = p.scale;
  double acc[kSelfAcc];
#pragma unroll
  for (int q = 0; q < kSelfAcc; ++q) acc[q] = 0.0;
  int qn = 0;
  // product modes: PQ and QQ are not written; this particle's rows of
  // PQ ein and QQ ein (prod), or its rows of PQ^T xt (tmode), are formed
  // from the pairs instead (pair parts unscaled here, per lane).  pr holds,
  // prod: PQ ein rows 3 gm + k (pr[k]) and 3 N + 3 gm + k (pr[3 + k]), QQ ein
  // rows 5 gm + I (pr[6 + I]); tmode: PQ^T xt rows 5 gm + I (pr[I]).
  constexpr bool prod = MODE == kLubProducts, tmode = MODE == kLubTranspose, bmode = MODE == kLubBsr;
  int pdone = 0;  // pairs of this particle processed in earlier batches (block-sparse slot base)
  double pr[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) pr[q] = 0.0;

  auto process = [&](int n) {
    // one queued pair per lane
    bool cross = false;
    if (lane < n) {
      const int dl = q_idx[lane];
      const int lo = gm < dl ? gm : dl;
      const double rx = q_geo[lane][0], ry = q_geo[lane][1], rz = q_geo[lane][2], s = q_geo[lane][3];
      const double r = sqrt(rx * rx + ry * ry + rz * rz);
      double v[19];
      lubcalc(s, *tab, v);
      const double e[3] = {rx / r, ry / r, rz / r};
      YProj yp;
      yproj(e[0], e[1], e[2], yp);
      // self contribution of gamma: (1,1) block if gamma is the lower index, else (2,2)
      const double sg = (gm == lo) ? 1.0 : -1.0;
      LubBlock b11;
      lub_block(e, v[0], v[2], v[4], v[6], v[8], v[10], v[12], v[14], yp, b11);
      // (the transpose mode uses only the g and h parts)
      double mm[15];
      if (!tmode) lub_mmx(v[16], v[17], v[18], yp, mm);
      if (!tmode) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          acc[q] += b11.a[q];
          acc[6 + q] += b11.c[q];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[12 + k] += sg * b11.b[k];
      }
#pragma unroll
      for (int q = 0; q < 15; ++q) {
        acc[15 + q] += sg * b11.g[q];
        acc[30 + q] += b11.h[q];
        if (!tmode) acc[45 + q] += mm[q];
      }
      LubBlock b12;
      if (prod || tmode || bmode || gm == lo)
        lub_block(e, v[1], v[3], v[5], v[7], v[9], v[11], v[13], v[15], yp, b12);
      if (bmode) {
        const int k = pdone + lane;
        if (k < p.bsr_cap) {
          double w[kCrossVals];
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            w[q] = sc * b12.a[q];
            w[6 + q] = sc * b12.c[q];
          }
#pragma unroll
          for (int q = 0; q < 3; ++q) w[12 + q] = sc * b12.b[q];
#pragma unroll
          for (int q = 0; q < 15; ++q) {
            w[15 + q] = sc * b12.g[q];
            w[30 + q] = sc * b12.h[q];
            w[45 + q] = sc * mm[q];
          }
          bsr_cross_block(w, gm == lo ? 1.0 : -1.0, p.bsr_blk + 121 * ((size_t)gm * p.bsr_cap + k));
          p.bsr_col[(size_t)gm * p.bsr_cap + k] = dl;
          if (p.bsr_dist) p.bsr_dist[(size_t)gm * p.bsr_cap + k] = r;
        } else {
          atomicOr(p.bsr_overflow, 1);
        }
      }
      if (prod || tmode) {
        // the pair's PQ entries in this particle's rows and columns, placed as
        // hydro_lub.f90:257-265 and :305 place them: g enters as -g where this
        // particle is the lower index, h as it is
        const double sgn = (gm == lo) ? -1.0 : 1.0;
        if (prod) {
          double e5[5];
#pragma unroll
          for (int I = 0; I < 5; ++I) e5[I] = p.ein[5 * dl + I];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double tg = 0.0, th = 0.0;
#pragma unroll
            for (int I = 0; I < 5; ++I) {
              tg = fma(b12.g[I + 5 * k], e5[I], tg);
              th = fma(b12.h[I + 5 * k], e5[I], th);
            }
            pr[k] = fma(sgn, tg, pr[k]);
            pr[3 + k] += th;
          }
#pragma unroll
          for (int I = 0; I < 5; ++I) {
            double tq = 0.0;
#pragma unroll
            for (int J = 0; J < 5; ++J) tq = fma(mm[packed5(I, J)], e5[J], tq);
            pr[6 + I] += tq;
          }
        } else {
          double xT[3], xR[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            xT[k] = p.xt[3 * dl + k];
            xR[k] = p.xt[3 * N + 3 * dl + k];
          }
#pragma unroll
          for (int I = 0; I < 5; ++I) {
            double tg = 0.0, th = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              tg = fma(b12.g[I + 5 * k], xT[k], tg);
              th = fma(b12.h[I + 5 * k], xR[k], th);
            }
            pr[I] += fma(-sgn, tg, th);
          }
        }
      }
      if (gm == lo && !tmode && !bmode) {
        cross = true;
        double* st = stage[lane];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          st[q] = sc * b12.a[q];
          st[6 + q] = sc * b12.c[q];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) st[12 + k] = sc * b12.b[k];
        if (!prod) {  // the products mode scatters only the PP entries (a, c, b)
#pragma unroll
          for (int q = 0; q < 15; ++q) {
            st[15 + q] = sc * b12.g[q];
            st[30 + q] = sc * b12.h[q];
            st[45 + q] = sc * mm[q];
          }
        }
        stage_be[lane] = dl;
      }
    }
    unsigned long long cm = __ballot(cross);
    pdone += n;
    __syncthreads();
    // cross blocks of the staged pairs, kScatterPairs pairs at a time: each
    // lane owns whole matrix entries (three per lane and pair), and all the
    // loads of a group go out before its stores, one memory round trip per
    // group (the entries of different pairs are distinct)
    while (cm) {
      int kq[kScatterPairs];
#pragma unroll
      for (int g = 0; g < kScatterPairs; ++g) {
        kq[g] = cm ? __ffsll((long long)cm) - 1 : -1;
        cm &= cm - 1;
      }
      double val[kScatterPairs][3], old[kScatterPairs][3];
      double* ptr[kScatterPairs][3];
#pragma unroll
      for (int g = 0; g < kScatterPairs; ++g)
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          const int e = lane + 64 * u;
          ptr[g][u] = kq[g] >= 0 && e < (prod ? kCrossPPEntries : kCrossEntries)
                          ? cross_entry(p, stage[kq[g]], gm, stage_be[kq[g]], e, val[g][u])
                          : nullptr;
        }
#pragma unroll
      for (int g = 0; g < kScatterPairs; ++g)
#pragma unroll
        for (int u = 0; u < 3; ++u)
          old[g][u] = ptr[g][u] && !(p.pq_zero && lane + 64 * u >= kCrossPPEntries) ? *ptr[g][u] : 0.0;
#pragma unroll
      for (int g = 0; g < kScatterPairs; ++g)
#pragma unroll
        for (int u = 0; u < 3; ++u)
          if (ptr[g][u]) *ptr[g][u] = old[g][u] + val[g][u];
    }
    __syncthreads();
  };

  for (int base = 0; base < N; base += 64) {
    const int dl = base + lane;
    bool take = false;
    double rx = 0.0, ry = 0.0, rz = 0.0, s = 0.0;
    if (dl < N && dl != gm) {
      const int lo = gm < dl ? gm : dl, hi = gm < dl ? dl : gm;
      if (lo <= N - p.Nb - 1) {  // hydro_lub.f90:68-71
        rx = pos[3 * lo] - pos[3 * hi];
        ry = pos[3 * lo + 1] - pos[3 * hi + 1];
        rz = pos[3 * lo + 2] - pos[3 * hi + 2];
        double r = sqrt(rx * rx + ry * ry + rz * rz);
        if (p.is_period) r = lub_skew(p, rx, ry, rz);
        const double r_ave = 0.5 * (radii[lo] + radii[hi]);
        s = r / r_ave;
        if (s <= 2.0) {
          s = 4.00001 - s;  // contact reflection, hydro_lub.f90:85-92
          if (gm == lo && contacts) atomicAdd(contacts, 1);
          take = true;
        } else {
          take = s <= 4.0;
        }
      }
    }
    const unsigned long long tm = __ballot(take);
    if (take) {
      const int slot = qn + __popcll(tm & ((1ull << lane) - 1ull));
      q_idx[slot] = dl;
      q_geo[slot][0] = rx;
      q_geo[slot][1] = ry;
      q_geo[slot][2] = rz;
      q_geo[slot][3] = s;
    }
    qn += __popcll(tm);
    __syncthreads();
    if (qn >= 64) {
      process(64);
      // move the rest (fewer than 64) to the front of the queue
      const int rest = qn - 64;
      int ri = 0;
      double rg[4] = {0.0, 0.0, 0.0, 0.0};
      if (lane < rest) {
        ri = q_idx[64 + lane];
#pragma unroll
        for (int c = 0; c < 4; ++c) rg[c] = q_geo[64 + lane][c];
      }
      __syncthreads();
      if (lane < rest) {
        q_idx[lane] = ri;
#pragma unroll
        for (int c = 0; c < 4; ++c) q_geo[lane][c] = rg[c];
      }
      __syncthreads();
      qn = rest;
    }
  }
  if (qn > 0) process(qn);

  // fixed-order wave reduction of the self accumulators, then the self blocks
  double* sv = stage[0];
#pragma unroll
  for (int q = tmode ? 15 : 0; q < (tmode ? 45 : kSelfAcc); ++q) {
    double x = acc[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) sv[q] = sc * x;
  }
  if (prod || tmode) {
#pragma unroll
    for (int q = 0; q < (tmode ? 5 : 11); ++q)
      for (int off = 32; off > 0; off >>= 1) pr[q] += __shfl_xor(pr[q], off, 64);
  }
  __syncthreads();
  if (prod && lane == 0) {
    double e5[5];
#pragma unroll
    for (int I = 0; I < 5; ++I) e5[I] = p.ein[5 * gm + I];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double tg = 0.0, th = 0.0;
#pragma unroll
      for (int I = 0; I < 5; ++I) {
        tg = fma(sv[15 + I + 5 * k], e5[I], tg);
        th = fma(sv[30 + I + 5 * k], e5[I], th);
      }
      p.pq_ein[3 * gm + k] = fma(sc, pr[k], tg);
      p.pq_ein[3 * N + 3 * gm + k] = fma(sc, pr[3 + k], th);
    }
#pragma unroll
    for (int I = 0; I < 5; ++I) {
      double tq = 0.0;
#pragma unroll
      for (int J = 0; J < 5; ++J) tq = fma(sv[45 + packed5(I, J)], e5[J], tq);
      p.qq_ein[5 * gm + I] = fma(sc, pr[6 + I], tq);
    }
  }
  if (bmode) {
    if (lane == 0) {
      bsr_self_block(sv, p.bsr_self + 121 * (size_t)gm);
      p.bsr_nnb[gm] = pdone < p.bsr_cap ? pdone : p.bsr_cap;
    }
    return;
  }
  if (tmode) {
    // nothing is written to the matrices in this mode
    if (lane == 0) {
      double xT[3], xR[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        xT[k] = p.xt[3 * gm + k];
        xR[k] = p.xt[3 * N + 3 * gm + k];
      }
#pragma unroll
      for (int I = 0; I < 5; ++I) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) t = fma(sv[15 + I + 5 * k], xT[k], fma(sv[30 + I + 5 * k], xR[k], t));
        p.pqt_x[5 * gm + I] = fma(sc, pr[I], t);
      }
    }
    return;
  }
  double val[2], old[2];
  double* ptr[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = lane + 64 * u;
    ptr[u] = e < (prod ? kSelfPPEntries : kSelfEntries) ? self_entry(p, sv, gm, e, val[u]) : nullptr;
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) old[u] = ptr[u] && !(p.pq_zero && lane + 64 * u >= kSelfPPEntries) ? *ptr[u] : 0.0;
#pragma unroll
  for (int u = 0; u < 2; ++u)
    if (ptr[u]) *ptr[u] = old[u] + val[u];
}

int lub_add(const LubParams& p, const double* d_pos, const double* d_radii, const LubTables* d_tab, int* d_contacts,
            hipStream_t s) {
  if (p.N < 1) return RSD_OK;
  hipEvent_t pe = prof_begin(kProfLub, 0.0, 0.0, s);
  if (p.bsr_blk)
    hipLaunchKernelGGL(lub_kernel<kLubBsr>, dim3(p.N), dim3(64), 0, s, p, d_pos, d_radii, d_tab, d_contacts);
  else if (p.xt)
    hipLaunchKernelGGL(lub_kernel<kLubTranspose>, dim3(p.N), dim3(64), 0, s, p, d_pos, d_radii, d_tab, d_contacts);
  else if (p.ein)
    hipLaunchKernelGGL(lub_kernel<kLubProducts>, dim3(p.N), dim3(64), 0, s, p, d_pos, d_radii, d_tab, d_contacts);
  else
    hipLaunchKernelGGL(lub_kernel<kLubDense>, dim3(p.N), dim3(64), 0, s, p, d_pos, d_radii, d_tab, d_contacts);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

}  // namespace rsd
