// rsd_internal.h -- shared declarations of the MI355X regularized Stokesian
// dynamics library (device kernels are in rsd_*.hip; the C ABI is rsd_capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rsd.h"

namespace rsd {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr int kTile = 64;  // dense tile edge; every dense matrix is padded to a multiple

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define RSD_HIP(call)                                                   \
  do {                                                                  \
    hipError_t _e = (call);                                             \
    if (_e != hipSuccess) return ::rsd::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define RSD_TRY(x)                 \
  do {                             \
    int _rc = (x);                 \
    if (_rc != RSD_OK) return _rc; \
  } while (0)

// ------------------------------------------------------------------ dense
// All dense routines act on column-major FP64 matrices whose order n is a
// multiple of kTile; callers pad with an identity block.
struct GemmDesc {
  const double* A;
  const double* B;
  double* C;
  int lda, ldb, ldc;
  int mt, nt, kt;  // C is (mt*64) x (nt*64), inner dimension kt*64
  double alpha, beta;
  int tile_base;   // first linear tile index of this descriptor in the launch
  int pad;
};

enum GemmFlags : int {
  kKloRow = 1,   // inner range starts at the output tile's row  (op(A) upper-triangular)
  kKloCol = 2,   // inner range starts at the output tile's column (op(B) lower-triangular)
  kKhiRow = 4,   // inner range ends after the output tile's row  (op(A) lower-triangular)
  kLowerC = 8,   // only output tiles on or below the diagonal (mt == nt)
};

struct DenseWork {
  // scratch owned by a context, grown on demand
  double* dinv = nullptr;   // n/64 inverted 64x64 diagonal factors
  size_t dinv_cap = 0;
  double* tmp = nullptr;    // n x n scratch for the triangular inverse
  size_t tmp_cap = 0;
  GemmDesc* d_desc = nullptr;  // device descriptor buffer
  size_t desc_cap = 0;
  int* d_info = nullptr;
  double* gv_part = nullptr;  // GEMV partial sums
  size_t gv_cap = 0;
  // triangular-solve hand-off: [0..1] tickets, [2] error word (a bounded wait
  // that ran out), then two sets of per-block ready flags holding the
  // generation of the solve that set them
  int* d_sync = nullptr;
  size_t sync_cap = 0;
  int gen = 0;
  // dataflow Cholesky hand-off: [0] ticket, [1] error word, nt*nt tile
  // flags, 2 nt partial-tile flags; and the partial tiles handed to the walker
  int* d_dag = nullptr;
  size_t dag_cap = 0;
  int* d_rec = nullptr;  // the dataflow Cholesky's first timed-out wait
  int dag_gen = 0;
  double* d_part = nullptr;
  size_t part_cap = 0;
};

int dense_reserve(DenseWork& w, int n, hipStream_t s);
void dense_release(DenseWork& w);
// C = alpha*op(A)*op(B) + beta*C for a batch of descriptors (host array)
int gemm_batched(DenseWork& w, const std::vector<GemmDesc>& descs, bool ta, bool tb, int flags, hipStream_t s);
// sets the device INFO word to "no failure"; potrf_lower keeps the minimum
// failing order over all factorizations since the last reset
int dense_info_reset(DenseWork& w, hipStream_t s);
// lower Cholesky in place; *info (device) gets the LAPACK-style INFO (min over failures)
int potrf_lower(DenseWork& w, double* A, int n, int ld, hipStream_t s);
// potrf_lower that also forms X = inv(L_S) for the trailing block L_S of the
// factor (tiles s0..n/64-1) in the same launch: X's lower blocks (and its
// diagonal blocks, zero above the diagonal) at X, leading dimension ldx; its
// upper blocks are overwritten with the transposes of the lower ones.
// s0 < 0 or X null: plain potrf_lower.
int potrf_lower_x(DenseWork& w, double* A, int n, int ld, int s0, double* X, int ldx, hipStream_t s);
// X(n x nrhs, leading ldx) := (L L^T)^-1 X with the factor potrf_lower just
// left in A (and its diagonal inverses in w); nrhs <= 2
int potrs_lower(DenseWork& w, const double* A, int n, int ld, double* X, int ldx, int nrhs, hipStream_t s);
// nonzero if a triangular solve since the last reset ran out of its bounded wait
int dense_sync_error(DenseWork& w, hipStream_t s, int* err);
// device addresses of the two bounded-wait error words (triangular solves,
// dataflow factorisation), or null before the first launch of each
const int* dense_solve_error_word(const DenseWork& w);
const int* dense_factor_error_word(const DenseWork& w);
// The dataflow Cholesky's first timed-out wait since dense_wait_reset: [1,
// waiting ticket, flag index (i + nt k for tile (i,k); nt^2 + p for partial
// tile p), the flag's value], or zeros.  dense_wait_reset clears it and the
// factorisation's error word (stream-ordered).
int dense_factor_wait_record(const DenseWork& w, int rec[4]);
int dense_wait_reset(DenseWork& w, hipStream_t s);
// in-place inverse of the lower factor using its 64x64 diagonal inverses
// (dinv, default: the ones potrf_lower just left in w)
int trtri_lower(DenseWork& w, double* A, int n, int ld, hipStream_t s, const double* dinv = nullptr);
// x := inv(L) x (trans: inv(L)^T x) for a lower factor with the given 64x64
// diagonal inverses; x has n entries (one persistent launch)
int trsv_lower(DenseWork& w, const double* L, int n, int ld, const double* dinv, double* x, bool trans,
               hipStream_t s);
// the padding rows and columns of the reordered grand mobility [C B^T; B A]
// (C's n5..n5p-1, A's n5p+n6..n5p+n6p-1) set to the identity's
int grand_padding(double* Gp, int ld, int n5, int n5p, int n6, int n6p, hipStream_t s);
// C(lower) = X^T X for lower-triangular X (LAPACK dlauum semantics, out of place)
int lauum_lower(DenseWork& w, const double* X, int ldx, double* C, int ldc, int n, hipStream_t s);
// The CG preconditioner in FP32 (n a multiple of 64): an FP32 copy of a
// potrf_lower factor's strictly lower tiles with w.dinv's diagonal-tile
// inverses; trtri_lower's recursion in FP32 on it (T: n x n FP32 scratch);
// C = X^T X for the lower tiles; the upper triangle from the lower.
int factor_to_f32(DenseWork& w, const double* L, int n, int ld, float* F, int ldf, hipStream_t s);
int trtri_lower32(DenseWork& w, float* A, int n, int ld, float* T, hipStream_t s);
int lauum_lower32(DenseWork& w, const float* X, int ldx, float* C, int ldc, int n, hipStream_t s);
int sym_fill_upper_f32(float* P, int n, hipStream_t s);
// zero the strict upper triangle
int zero_strict_upper(double* A, int n, int ld, hipStream_t s);
// fill the strict upper triangle from the lower one
int symmetrize_from_lower(double* A, int n, int ld, hipStream_t s);
// A(n x n, leading ld) := sym(upper of src) with identity padding up to np
int load_sym_from_upper(const double* src, int n, int ldsrc, double* A, int np, int ld, hipStream_t s);
// SPD inverse: A (np x np padded) -> Ainv (np x np), both leading dim ld; A is overwritten
int spd_inverse(DenseWork& w, double* A, double* Ainv, int np, int ld, hipStream_t s);
// y = alpha*op(A) x + beta*y  (A m x n column-major)
int gemv(DenseWork& w, const double* A, int lda, int m, int n, bool trans, double alpha, const double* x, double beta,
         double* y, hipStream_t s);
int set_identity_padding(double* A, int n, int np, int ld, hipStream_t s);
int copy_block(const double* src, int lds, double* dst, int ldd, int m, int n, hipStream_t s);
// Y(m x n) += alpha * X(m x n)
int matrix_axpy(double alpha, const double* X, int ldx, double* Y, int ldy, int m, int n, hipStream_t s);
// y = a + b (n entries; y may not alias a or b)
int vector_sum(const double* a, const double* b, double* y, int n, hipStream_t s);

}  // namespace rsd
