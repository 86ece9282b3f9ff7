// rsd_pcg.h -- preconditioned conjugate gradients for the collision sub-step
// solves (rsd_pcg.hip).
#pragma once
#include "rsd_internal.h"

namespace rsd {

enum PcgStatus : int {
  kPcgRunning = 0,
  kPcgConverged = 1,
  kPcgBreakdown = 2,     // p.Ap <= 0 or NaN: the operator is not SPD (as far as CG can tell)
  kPcgNotConverged = 3,  // max_iter reached
};

constexpr int kPcgMaxSlots = 256;  // solves whose status one step can hold (sub-steps x right-hand sides)

struct PcgWork {
  float* P = nullptr;       // n x n FP32 inverse of an earlier operator (symmetric, full)
  double* vec = nullptr;    // r[2], z, p[2], q, spare: 7 n
  double* part = nullptr;   // per-workgroup partial sums
  int* state = nullptr;     // per solve slot: [0] PcgStatus, [1] iterations (device address of h_state)
  unsigned* bar = nullptr;  // [0] grid-barrier arrivals, [1] barrier-wait error word
  int* h_state = nullptr;   // the slots, in pinned host memory the kernels write
  int n_cap = 0;
  int n = 0;                // order of the current preconditioner (0: none)
  int persist_G = 0;        // workgroups of the one-launch solver (0: not available)
  int cus = 0;              // compute units of the device (workgroups of the phased one-per-CU solver)
  int part_blk = 0;         // partial-sum entries per slot that part holds
  bool probed = false;
};

int pcg_reserve(PcgWork& pw, int n);
void pcg_release(PcgWork& pw);
// P := inv(L L^T), computed in FP32 from the lower factor potrf_lower just
// left in L (and its diagonal inverses in w): an FP32 copy of the factor in
// w's scratch, its triangular inverse, the Gram product.  L is only read, by
// the first launch; with copied, that event is recorded on s once L is read.
int pcg_precond_from_factor(DenseWork& w, PcgWork& pw, const double* L, int n, int ld, hipStream_t s,
                            hipEvent_t copied = nullptr);
// Enqueues the solve of A x = b (A symmetric, n x n, leading lda, full
// storage) from the initial guess in x, preconditioned by P, until
// ||b - A x|| <= tol ||b|| or max_iter iterations.  Its PcgStatus and
// iteration count land in status slot `slot` on the device.  With one_launch,
// for n up to 8192 and a device that runs cooperative launches, the whole
// solve is one launch and the host does not wait; otherwise each phase is a
// launch and the host checks the status after chunk_hint iterations, then
// every few.
int pcg_solve(PcgWork& pw, const double* A, int n, int lda, const double* b, double* x, double tol, int max_iter,
              int slot, bool one_launch, int chunk_hint, hipStream_t s);
// Device address of the one-launch solver's barrier-wait error word (set
// when a wait ran out; whoever reads it clears it), or null before the first
// solve.  The status slots [0, nslots) (status, iterations pairs) are written
// by the kernels into pw.h_state itself, valid once the stream has synchronised.
unsigned* pcg_error_word(PcgWork& pw);
// The CG initial guess of a sub-step from the two before it: x (the last
// solution) := 2 x - xh and xh := the last solution; with first (one
// solution so far) only xh := x.
int pcg_initial_guess(double* x, double* xh, int n, bool first, hipStream_t s);
// true when solves of order n run as one launch
bool pcg_one_launch(const PcgWork& pw, int n);

}  // namespace rsd
