// rsd_lub.h -- launch interface of the lubrication kernel.
#pragma once
#include "rsd_internal.h"
#include "rsd_lubmath.h"

namespace rsd {

struct LubParams {
  double LR[9], LI[9];  // lattice used by the minimum image
  double scale;         // 6 pi mu_f (hydro_lub.f90:304-306)
  int N, Nb, is_period;
  int pq_zero;          // PQ and QQ hold zeros wherever the lubrication goes: stored, not added to
  double* PP;           // rfu target (6N x 6N block), added to
  double* PQ;           // rfe target (6N x 5N block)
  double* QQ;           // rse target (5N x 5N block)
  int ldPP, ldPQ, ldQQ, pad2;
  // Product modes (both leave PQ and QQ untouched): with ein (5N), this
  // particle's rows of PQ ein go to pq_ein (6N) and of QQ ein to qq_ein (5N),
  // PP is added to as usual; with xt (6N), only PQ^T xt goes to pqt_x (5N)
  // and nothing else is written.
  const double* ein = nullptr;
  double* pq_ein = nullptr;
  double* qq_ein = nullptr;
  const double* xt = nullptr;
  double* pqt_x = nullptr;
  // Block-sparse mode (bsr_blk set): nothing is added to PP/PQ/QQ; instead
  // each particle g gets its row of the grand lubrication resistance
  // [rfu rfe; rfe^T rse] as 11x11 blocks in its own dof order (3 T, 3 R, 5 Q),
  // column-major: the self block at bsr_self + 121 g, and its neighbours'
  // blocks at bsr_blk + 121 (g cap + k), k < bsr_nnb[g], neighbour index
  // bsr_col[g cap + k] in increasing order; *bsr_overflow is set when a
  // particle has more than cap neighbours in range.
  double* bsr_blk = nullptr;
  double* bsr_self = nullptr;
  int* bsr_col = nullptr;
  int* bsr_nnb = nullptr;
  int bsr_cap = 0;
  int* bsr_overflow = nullptr;
  double* bsr_dist = nullptr;  // optional: the pair's (minimum-image) centre distance per slot
};

// Adds the lubrication resistance into the three targets (which the caller
// zeroes for the stand-alone lubmxcalc, or fills with M^-1 for the step).
int lub_add(const LubParams& p, const double* d_pos, const double* d_radii, const rsdm::LubTables* d_tab,
            int* d_contacts, hipStream_t s);

}  // namespace rsd
