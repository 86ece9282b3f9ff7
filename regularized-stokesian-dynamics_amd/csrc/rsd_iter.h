// rsd_iter.h -- the iterative resistance solve (rsd_iter.hip): flexible GMRES
// on the mobility form of the grand resistance problem, the form the
// reference's hydro_SD_FTS (stepper.f90:476-546) solves with GMRES
// (gmres_mod_gmres.f90:23-61) instead of inverting the grand mobility M.
//
// With R = inv(M) + R_lub and the far-field force x = inv(M) [U; -EIN], the
// known force F and the known strain EIN fix
//     K x = [F; -EIN],   K = [ (I + R_lub M)_P ; M_Q ]   (11N x 11N),
// and then U = (M x)_P, SijN = (x + R_lub M x)_Q.  A product with K is one
// dense product with M (HBM-bound) and one block-sparse product with the
// lubrication (11x11 blocks of the neighbour pairs).  Vectors are
// field-major (T 3N, R 3N, Q 5N) as the reference orders them.
#pragma once
#include "rsd_internal.h"

namespace rsd {

// The lubrication resistance in 11x11 blocks (LubParams::bsr_*), in the
// precision T of its blocks.
template <typename T>
struct LubBsr {
  int N = 0, cap = 0;
  const int* nnb = nullptr;
  const int* col = nullptr;
  const T* blk = nullptr;
  const T* self = nullptr;
};

enum IterPrecond : int {
  kIterNone = 0,         // unpreconditioned (the reference's GMRES)
  kIterBlockJacobi = 1,  // inverses of K's 11x11 diagonal blocks
  kIterMixed = 2,        // an inner FP32 GMRES on FP32 copies of M and R_lub (block-Jacobi inside)
  kIterLearned = 3,      // block-Jacobi plus a learned block-sparse correction
  kIterMixedLearned = 4, // the FP32 inner GMRES with the learned preconditioner inside
};

struct IterWork {
  int n = 0, m = 0;          // order 11N, restart length
  double* V = nullptr;       // (m+1) Krylov vectors
  double* Z = nullptr;       // m preconditioned vectors (flexible GMRES)
  double* w = nullptr;       // work: 4 vectors
  double* part = nullptr;    // split-K partial sums
  double* h = nullptr;       // device Hessenberg column + norm
  double* hh = nullptr;      // pinned host copy
  float* V32 = nullptr;      // FP32 inner solver's Krylov and work vectors
  float* w32 = nullptr;
  float* part32 = nullptr;
  float* h32 = nullptr;
  float* hh32 = nullptr;
  double* dinv = nullptr;    // block-Jacobi inverses, 121 per particle
  float* dinv32 = nullptr;
  double* corr = nullptr;    // learned correction blocks (self + neighbour slots), 121 each
  size_t cap_n = 0;
  // the FP32 inner GMRES captured as a HIP graph on fixed buffers (vin -> zout),
  // re-captured when what it was captured with changes
  double* vin = nullptr;
  double* zout = nullptr;
  hipGraph_t inner_graph = nullptr;
  hipGraphExec_t inner_exec = nullptr;
  const void* inner_key[8] = {};
  int inner_key_i[4] = {};
};

int iter_reserve(IterWork& w, int n11, int m, int N, int bsr_cap);
void iter_release(IterWork& w);

// y = M x for a dense column-major M (n x n, leading ld), split over K.
template <typename T>
int dense_gemv(const T* M, int ld, int n, const T* x, T* y, T* part, hipStream_t s);
// y = x_P + (R_lub (M x))_P rows, y_Q = (M x)_Q: the operator K.
template <typename T>
int apply_k(const LubBsr<T>& L, const T* M, int ld, const T* x, T* y, T* wbuf, T* part, hipStream_t s);
// lw = R_lub w (all 11N rows)
template <typename T>
int bsr_apply(const LubBsr<T>& L, const T* w, T* lw, bool p_rows_only, hipStream_t s);

// The inverses of K's 11x11 diagonal blocks (particle dof order), FP64:
// rows P of block g: I + sum_d R_lub(g,d) M(d,g); rows Q: M(g,g)_Q.
int block_jacobi_build(const LubBsr<double>& L, const double* M, int ld, double* dinv, hipStream_t s);
// z = blockdiag(dinv) v (+ the neighbour blocks of corr when given)
template <typename T>
int block_apply(int N, const T* dinv, const double* corr, const LubBsr<double>* Lcorr, const T* v, T* z,
                hipStream_t s);

struct GmresResult {
  int iterations = 0;
  double relres = 0.0;  // final relative residual ||b - K x|| / ||b|| (recomputed)
  bool converged = false;
  bool at_floor = false;  // stalled at FP64's floor within 100 tol, taken as converged
};

// x := solution of K x = b by right-preconditioned flexible GMRES(m) in FP64,
// from the x given, until ||b - K x|| <= tol ||b|| or max_iter iterations, or
// until a restart cycle stalls at FP64's floor (within 100 tol: converged).
// precond kIterMixed runs inner_iters FP32 GMRES iterations (themselves
// block-Jacobi preconditioned, plus the learned correction with
// kIterMixedLearned) on K32 as the preconditioner.
struct GmresSetup {
  const LubBsr<double>* L;
  const double* M;
  int ld;
  const LubBsr<float>* L32;  // kIterMixed
  const float* M32;
  int ld32;
  int precond;
  int inner_iters;
  const LubBsr<double>* Lcorr;  // kIterLearned, kIterMixedLearned: the correction's neighbour structure
  bool graphs = true;           // the mixed preconditioners' inner GMRES replayed as a captured HIP graph
};
int gmres_solve(IterWork& w, const GmresSetup& g, const double* b, double* x, double tol, int max_iter,
                GmresResult* res, hipStream_t s);

// field-major <-> FP32 copies
int to_f32(const double* x, float* y, size_t n, hipStream_t s);
int to_f64(const float* x, double* y, size_t n, hipStream_t s);
int matrix_to_f32(const double* A, int lda, float* B, int ldb, int m, int n, hipStream_t s);

}  // namespace rsd
