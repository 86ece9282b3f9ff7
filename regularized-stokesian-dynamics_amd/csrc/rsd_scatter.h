// rsd_scatter.h -- device helpers shared by the grand-mobility kernels
// (Ewald under Lees-Edwards images, rsd_ewald.hip; free space, rsd_free.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "rsd_pairmath.h"

namespace rsd {

__device__ __forceinline__ void store(double* G, size_t ld, int r, int c, double v) { G[(size_t)r + (size_t)c * ld] = v; }

// Pair (alpha < beta) of lane `pid` when the N(N-1)/2 pairs are packed in
// row-major order (row alpha starts at alpha (N-1) - alpha (alpha-1)/2), so
// that every wave but the last is full.  Lanes past the last pair get the
// last pair and false.
__device__ __forceinline__ bool pair_of_lane(long long pid, int N, int& al, int& be) {
  const long long NP = N, np = NP * (NP - 1) / 2;
  const bool active = pid < np;
  const long long pq = active ? pid : np - 1;
  auto row_start = [NP](long long a) { return a * (NP - 1) - a * (a - 1) / 2; };
  const double tn = 2.0 * (double)NP - 1.0;
  long long a0 = (long long)((tn - sqrt(tn * tn - 8.0 * (double)pq)) * 0.5);
  a0 = a0 < 0 ? 0 : (a0 > NP - 2 ? NP - 2 : a0);
  while (a0 + 1 <= NP - 2 && row_start(a0 + 1) <= pq) ++a0;
  while (a0 > 0 && row_start(a0) > pq) --a0;
  al = (int)a0;
  be = al + 1 + (int)(pq - row_start(a0));
  return active;
}

// The 18 blocks of pair (al, be) into the grand mobility, scaled by sc
// (per_GRPY_YZD.f90:122-188, GRPY_YZD.f90:166-239): a and c copied to (be,al),
// b and g negated there, h and m copied; the transposed blocks by symmetry.
// oT, oR, oQ: rows (and columns) where particle 0's translation, rotation and
// stresslet fields start.
__device__ __forceinline__ void scatter_pair(double* __restrict__ G, size_t ld, int oT, int oR, int oQ, int al, int be,
                                             const rsdm::PairOut& o, double sc) {
  double TT[3][3], RR[3][3], TR[3][3];
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    TT[si[q]][sj[q]] = TT[sj[q]][si[q]] = o.tt[q] * sc;
    RR[si[q]][sj[q]] = RR[sj[q]][si[q]] = o.rr[q] * sc;
  }
  const double t0 = o.tr[0] * sc, t1 = o.tr[1] * sc, t2 = o.tr[2] * sc;
  TR[0][0] = 0.0;
  TR[0][1] = t2;
  TR[0][2] = -t1;
  TR[1][0] = -t2;
  TR[1][1] = 0.0;
  TR[1][2] = t0;
  TR[2][0] = t1;
  TR[2][1] = -t0;
  TR[2][2] = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    const int a = pass ? be : al, b = pass ? al : be;
    const double sg = pass ? -1.0 : 1.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        store(G, ld, oT + 3 * a + i, oT + 3 * b + j, TT[i][j]);
        store(G, ld, oT + 3 * a + i, oR + 3 * b + j, sg * TR[i][j]);
        store(G, ld, oR + 3 * a + i, oR + 3 * b + j, RR[i][j]);
        store(G, ld, oR + 3 * b + j, oT + 3 * a + i, sg * TR[i][j]);
      }
    for (int i = 0; i < 3; ++i)
      for (int J = 0; J < 5; ++J) {
        const double g = sg * o.gq[i + 3 * J] * sc, h = o.hq[i + 3 * J] * sc;
        store(G, ld, oT + 3 * a + i, oQ + 5 * b + J, g);
        store(G, ld, oR + 3 * a + i, oQ + 5 * b + J, h);
        store(G, ld, oQ + 5 * b + J, oT + 3 * a + i, g);
        store(G, ld, oQ + 5 * b + J, oR + 3 * a + i, h);
      }
    int k = 0;
    for (int I = 0; I < 5; ++I)
      for (int J = I; J < 5; ++J, ++k) {
        const double v = o.qq[k] * sc;
        store(G, ld, oQ + 5 * a + I, oQ + 5 * b + J, v);
        if (J != I) store(G, ld, oQ + 5 * a + J, oQ + 5 * b + I, v);
      }
  }
}

}  // namespace rsd
