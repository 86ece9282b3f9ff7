// rsd_iter.hip -- the iterative resistance solve: flexible GMRES on
//   K x = [F; -EIN],  K = [ (I + R_lub M)_P ; M_Q ],
// the mobility form the reference's hydro_SD_FTS (stepper.f90:476-546) hands
// to GMRES (gmres_mod_gmres.f90:23-61) instead of inverting M.
//
//   dense_gemv     y = M x, split over K: 256 rows x 64 columns a workgroup,
//                  the x chunk in LDS, a second pass sums the partials in a
//                  fixed order (HBM-bound: one pass over M)
//   bsr_apply      R_lub w from the 11x11 neighbour blocks, one thread per row
//   block_jacobi   K's 11x11 diagonal blocks, formed from the lubrication
//                  blocks and M, inverted by Gauss-Jordan with partial pivoting
//   gmres_solve    FP64 right-preconditioned flexible GMRES with classical
//                  Gram-Schmidt applied twice; the Hessenberg's Givens
//                  rotations on the host (one wait per iteration).  The mixed
//                  preconditioner is a fixed number of FP32 GMRES iterations
//                  on FP32 copies of M and R_lub, entirely on the device (its
//                  small least-squares problem solved by one thread), so the
//                  bulk of the matrix traffic moves half the bytes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "rsd_iter.h"
#include "rsd_prof.h"

namespace rsd {

constexpr int kGemvThreads = 128, kGemvCols = 64;

// particle g's local dof c (T 0-2, R 3-5, Q 6-10) in the field-major vector
__device__ __forceinline__ int dof(int N, int g, int c) {
  return c < 3 ? 3 * g + c : (c < 6 ? 3 * N + 3 * g + (c - 3) : 6 * N + 5 * g + (c - 6));
}

// 16-byte loads: each lane owns V consecutive rows (4 floats or 2 doubles), a
// workgroup kGemvThreads V rows and kGemvCols columns of M, x's chunk in LDS;
// the partial sums of the column chunk go to part (one row of partials per
// chunk).  Rows past the last full vector take scalar loads.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ void vec_fma(const typename Vec16<T>::type& v, T x, T* acc);
template <>
__device__ __forceinline__ void vec_fma<float>(const float4& v, float x, float* acc) {
  acc[0] = fmaf(v.x, x, acc[0]);
  acc[1] = fmaf(v.y, x, acc[1]);
  acc[2] = fmaf(v.z, x, acc[2]);
  acc[3] = fmaf(v.w, x, acc[3]);
}
template <>
__device__ __forceinline__ void vec_fma<double>(const double2& v, double x, double* acc) {
  acc[0] = fma(v.x, x, acc[0]);
  acc[1] = fma(v.y, x, acc[1]);
}

template <typename T>
__global__ __launch_bounds__(kGemvThreads) void gemv_split_kernel(const T* __restrict__ M, int ld, int n,
                                                                  const T* __restrict__ x, T* __restrict__ part) {
  constexpr int V = Vec16<T>::n;
  using VT = typename Vec16<T>::type;
  __shared__ T xs[kGemvCols];
  const int r0 = (blockIdx.x * kGemvThreads + threadIdx.x) * V, c0 = blockIdx.y * kGemvCols;
  const int nc = min(kGemvCols, n - c0);
  if (threadIdx.x < nc) xs[threadIdx.x] = x[c0 + threadIdx.x];
  __syncthreads();
  if (r0 >= n) return;
  T acc[V];
#pragma unroll
  for (int k = 0; k < V; ++k) acc[k] = 0;
  const T* col = M + (size_t)c0 * ld + r0;
  // the vector loads need ld and r0 multiples of V (16-byte aligned columns)
  if (r0 + V <= n && (ld % V) == 0) {
    if (nc == kGemvCols) {
#pragma unroll 16
      for (int c = 0; c < kGemvCols; ++c) vec_fma<T>(*reinterpret_cast<const VT*>(col + (size_t)c * ld), xs[c], acc);
    } else {
      for (int c = 0; c < nc; ++c) vec_fma<T>(*reinterpret_cast<const VT*>(col + (size_t)c * ld), xs[c], acc);
    }
  } else {
    for (int c = 0; c < nc; ++c)
#pragma unroll
      for (int k = 0; k < V; ++k)
        if (r0 + k < n) acc[k] = fma(col[(size_t)c * ld + k], xs[c], acc[k]);
  }
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (r0 + k < n) part[(size_t)blockIdx.y * n + r0 + k] = acc[k];
}

template <typename T>
__global__ void gemv_sum_kernel(int n, int splits, const T* __restrict__ part, T* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  T acc = 0;
  for (int k = 0; k < splits; ++k) acc += part[(size_t)k * n + r];
  y[r] = acc;
}

template <typename T>
int dense_gemv(const T* M, int ld, int n, const T* x, T* y, T* part, hipStream_t s) {
  constexpr int V = Vec16<T>::n;
  const int splits = (n + kGemvCols - 1) / kGemvCols;
  const int rows_per_block = kGemvThreads * V;
  // algorithmic work: 2 n^2 flops; bytes: M once, x, the partials written and read, y
  const double nn = (double)n * n, by = sizeof(T) * (nn + n + 2.0 * splits * n + n);
  hipEvent_t pe = prof_begin(sizeof(T) == 8 ? kProfGemvSplit64 : kProfGemvSplit32, 2.0 * nn, by, s);
  hipLaunchKernelGGL(gemv_split_kernel<T>, dim3((n + rows_per_block - 1) / rows_per_block, splits),
                     dim3(kGemvThreads), 0, s, M, ld, n, x, part);
  hipLaunchKernelGGL(gemv_sum_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, n, splits, (const T*)part, y);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// One wave per particle g: lane l takes the (slot, column) pairs l, l+64, ...
// of g's block row (slot 0 the self block, 1+q neighbour q), so that each
// lane reads whole 11-entry columns of the 11x11 blocks; the eleven row sums
// are then reduced across the wave.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// lw rows of particle g: self block and neighbour blocks times the
// neighbours' dofs (rows < 6 only when p_only)
template <typename T>
__global__ __launch_bounds__(64) void bsr_apply_kernel(LubBsr<T> L, const T* __restrict__ w, T* __restrict__ lw,
                                                       int nrows) {
  const int g = blockIdx.x, lane = threadIdx.x, N = L.N, nn = L.nnb[g];
  T acc[11];
#pragma unroll
  for (int r = 0; r < 11; ++r) acc[r] = 0;
  for (int it = lane; it < 11 * (nn + 1); it += 64) {
    const int slot = it / 11, c = it % 11;
    const int d = slot == 0 ? g : L.col[(size_t)g * L.cap + slot - 1];
    const T* b = slot == 0 ? L.self + 121 * (size_t)g : L.blk + 121 * ((size_t)g * L.cap + slot - 1);
    const T x = w[dof(N, d, c)];
#pragma unroll
    for (int r = 0; r < 11; ++r)
      if (r < nrows) acc[r] = fma(b[r + 11 * c], x, acc[r]);
  }
#pragma unroll
  for (int r = 0; r < 11; ++r) {
    const T v = wave_sum(acc[r]);
    if (lane == r && r < nrows) lw[dof(N, g, r)] = v;
  }
}

template <typename T>
int bsr_apply(const LubBsr<T>& L, const T* w, T* lw, bool p_rows_only, hipStream_t s) {
  const int nrows = p_rows_only ? 6 : 11;
  hipLaunchKernelGGL(bsr_apply_kernel<T>, dim3(L.N), dim3(64), 0, s, L, w, lw, nrows);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// y_P = x_P + lw_P, y_Q = w_Q
template <typename T>
__global__ void k_finish_kernel(int N, const T* __restrict__ x, const T* __restrict__ w, const T* __restrict__ lw,
                                T* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 11 * N) return;
  y[i] = i < 6 * N ? x[i] + lw[i] : w[i];
}

template <typename T>
int apply_k(const LubBsr<T>& L, const T* M, int ld, const T* x, T* y, T* wbuf, T* part, hipStream_t s) {
  const int n = 11 * L.N;
  T* w = wbuf;
  T* lw = wbuf + n;
  RSD_TRY(dense_gemv(M, ld, n, x, w, part, s));
  RSD_TRY(bsr_apply(L, (const T*)w, lw, true, s));
  hipLaunchKernelGGL(k_finish_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, L.N, x, (const T*)w,
                     (const T*)lw, y);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ----------------------------------------------------------- block Jacobi
__global__ __launch_bounds__(128) void block_jacobi_kernel(LubBsr<double> L, const double* __restrict__ M, int ld,
                                                           double* __restrict__ dinv) {
  __shared__ double A[11][12], X[11][12];
  __shared__ int piv;
  const int g = blockIdx.x, t = threadIdx.x, N = L.N;
  if (t < 121) {
    const int r = t % 11, c = t / 11;
    double v;
    if (r < 6) {
      v = (r == c) ? 1.0 : 0.0;
      // sum over the blocks of row g of R_lub (self and neighbours) of R(g,d)[r,:] M(d,g)[:,c]
      const double* b = L.self + 121 * (size_t)g;
      const int cg = dof(N, g, c);
      for (int k = 0; k < 11; ++k) v = fma(b[r + 11 * k], M[(size_t)dof(N, g, k) + (size_t)cg * ld], v);
      for (int q = 0; q < L.nnb[g]; ++q) {
        const int d = L.col[(size_t)g * L.cap + q];
        const double* bk = L.blk + 121 * ((size_t)g * L.cap + q);
        for (int k = 0; k < 11; ++k) v = fma(bk[r + 11 * k], M[(size_t)dof(N, d, k) + (size_t)cg * ld], v);
      }
    } else {
      v = M[(size_t)dof(N, g, r) + (size_t)dof(N, g, c) * ld];
    }
    A[r][c] = v;
    X[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  // Gauss-Jordan with partial pivoting: A -> I, X -> inv(A)
  for (int k = 0; k < 11; ++k) {
    if (t == 0) {
      int p = k;
      double best = fabs(A[k][k]);
      for (int i = k + 1; i < 11; ++i)
        if (fabs(A[i][k]) > best) {
          best = fabs(A[i][k]);
          p = i;
        }
      piv = p;
    }
    __syncthreads();
    if (piv != k && t < 11) {
      double a = A[k][t], b = A[piv][t];
      A[k][t] = b;
      A[piv][t] = a;
      a = X[k][t];
      b = X[piv][t];
      X[k][t] = b;
      X[piv][t] = a;
    }
    __syncthreads();
    const double inv = 1.0 / A[k][k];
    __syncthreads();
    if (t < 11) {
      A[k][t] *= inv;
      X[k][t] *= inv;
    }
    __syncthreads();
    double fa = 0.0, fx = 0.0, f = 0.0;
    int r = 0, c = 0;
    if (t < 121) {
      r = t % 11;
      c = t / 11;
      f = A[r][k];
      fa = A[k][c];
      fx = X[k][c];
    }
    __syncthreads();
    if (t < 121 && r != k) {
      A[r][c] -= f * fa;
      X[r][c] -= f * fx;
    }
    __syncthreads();
  }
  if (t < 121) dinv[121 * (size_t)g + (t % 11) + 11 * (t / 11)] = X[t % 11][t / 11];
}

int block_jacobi_build(const LubBsr<double>& L, const double* M, int ld, double* dinv, hipStream_t s) {
  hipLaunchKernelGGL(block_jacobi_kernel, dim3(L.N), dim3(128), 0, s, L, M, ld, dinv);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// z_g = dinv_g v_g + (corr_gg v_g + sum_d corr_gd v_d when a correction is
// given), one wave per particle as bsr_apply_kernel
template <typename T>
__global__ __launch_bounds__(64) void block_apply_kernel(int N, const T* __restrict__ dinv,
                                                         const double* __restrict__ corr, LubBsr<double> Lc,
                                                         const T* __restrict__ v, T* __restrict__ z) {
  const int g = blockIdx.x, lane = threadIdx.x;
  const int nn = corr ? Lc.nnb[g] : 0;
  const size_t base = (size_t)g * (Lc.cap + 1);
  T acc[11];
#pragma unroll
  for (int r = 0; r < 11; ++r) acc[r] = 0;
  // items 0..10: dinv's columns; 11..: the correction's (slot, column) pairs
  // (slot 0 its self block, 1+q neighbour q), only when there is a correction
  const int items = corr ? 11 * (nn + 2) : 11;
  for (int it = lane; it < items; it += 64) {
    if (it < 11) {
      const T x = v[dof(N, g, it)];
      const T* b = dinv + 121 * (size_t)g;
#pragma unroll
      for (int r = 0; r < 11; ++r) acc[r] = fma(b[r + 11 * it], x, acc[r]);
    } else {
      const int slot = it / 11 - 1, c = it % 11;
      const int d = slot == 0 ? g : Lc.col[(size_t)g * Lc.cap + slot - 1];
      const double* b = corr + 121 * (base + slot);
      const double x = (double)v[dof(N, d, c)];
#pragma unroll
      for (int r = 0; r < 11; ++r) acc[r] += (T)(b[r + 11 * c] * x);
    }
  }
#pragma unroll
  for (int r = 0; r < 11; ++r) {
    const T s = wave_sum(acc[r]);
    if (lane == r) z[dof(N, g, r)] = s;
  }
}

template <typename T>
int block_apply(int N, const T* dinv, const double* corr, const LubBsr<double>* Lcorr, const T* v, T* z,
                hipStream_t s) {
  LubBsr<double> Lc{};
  if (corr && Lcorr) Lc = *Lcorr;
  hipLaunchKernelGGL(block_apply_kernel<T>, dim3(N), dim3(64), 0, s, N, dinv, (corr && Lcorr) ? corr : nullptr, Lc,
                     v, z);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------------------------ vector kernels
// h[j] = V_j . w for j < k (one workgroup per j, fixed-order tree in LDS)
template <typename T>
__global__ __launch_bounds__(256) void multi_dot_kernel(const T* __restrict__ V, size_t ldv, int n,
                                                        const T* __restrict__ w, T* __restrict__ h) {
  __shared__ double red[256];
  const T* v = V + (size_t)blockIdx.x * ldv;
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) a = fma((double)v[i], (double)w[i], a);
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) h[blockIdx.x] = (T)red[0];
}

// w -= V h (k columns)
template <typename T>
__global__ void sub_vh_kernel(const T* __restrict__ V, size_t ldv, int n, int k, const T* __restrict__ h,
                              T* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T a = w[i];
  for (int j = 0; j < k; ++j) a -= V[(size_t)j * ldv + i] * h[j];
  w[i] = a;
}

// x += Z y (k columns)
template <typename T>
__global__ void add_zy_kernel(const T* __restrict__ Z, size_t ldz, int n, int k, const T* __restrict__ y,
                              T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T a = x[i];
  for (int j = 0; j < k; ++j) a += Z[(size_t)j * ldz + i] * y[j];
  x[i] = a;
}

// y = x * s where s = scal_num / *scal_den (device scalars; den null: 1)
template <typename T>
__global__ void scale_kernel(int n, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ den, T num,
                             int recip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T d = den ? *den : (T)1;
  y[i] = recip ? x[i] * (num / d) : x[i] * (num * d);
}

// x += y
template <typename T>
__global__ void vec_add_kernel(int n, const T* __restrict__ y, T* __restrict__ x) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] += y[i];
}

// r = b - y
__global__ void residual_kernel(int n, const double* __restrict__ b, const double* __restrict__ y,
                                double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

// h[k] = sqrt(h[k]) (the last entry of a multi_dot of w with itself)
template <typename T>
__global__ void sqrt_kernel(T* h) {
  if (threadIdx.x == 0) h[0] = sqrt(h[0]);
}

__global__ void f64_to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
__global__ void f32_to_f64_kernel(const float* __restrict__ x, double* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}
__global__ void matrix_to_f32_kernel(const double* __restrict__ A, int lda, float* __restrict__ B, int ldb, int m,
                                     int n) {
  const int c = blockIdx.y;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x)
    B[(size_t)r + (size_t)c * ldb] = (float)A[(size_t)r + (size_t)c * lda];
}

int to_f32(const double* x, float* y, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, x,
                     y, n);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}
int to_f64(const float* x, double* y, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, x,
                     y, n);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}
int matrix_to_f32(const double* A, int lda, float* B, int ldb, int m, int n, hipStream_t s) {
  hipLaunchKernelGGL(matrix_to_f32_kernel, dim3((m + 255) / 256, n), dim3(256), 0, s, A, lda, B, ldb, m, n);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------------ the FP32 inner GMRES (device only)
// Least squares of the (k+1) x k Hessenberg H (column j at H + j*ldh) with
// right-hand side beta e1 by Givens rotations, in FP64 in LDS; y[0..k) out.
// k <= 31; one thread does the arithmetic.
__global__ __launch_bounds__(64) void inner_lsq_kernel(const float* __restrict__ H, int ldh, int k,
                                                       const float* __restrict__ beta, float* __restrict__ y) {
  __shared__ double R[32][33];
  __shared__ double g[33], cs[32], sn[32], yd[32];
  for (int e = threadIdx.x; e < 32 * 32; e += 64) {
    const int i = e % 32, j = e / 32;
    R[i][j] = (j < k && i <= j + 1) ? (double)H[(size_t)j * ldh + i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i <= k; ++i) g[i] = 0.0;
  g[0] = beta[0];
  for (int j = 0; j < k; ++j) {
    for (int i = 0; i < j; ++i) {
      const double a = R[i][j], b = R[i + 1][j];
      R[i][j] = cs[i] * a + sn[i] * b;
      R[i + 1][j] = -sn[i] * a + cs[i] * b;
    }
    const double a = R[j][j], b = R[j + 1][j], rr = sqrt(a * a + b * b);
    cs[j] = rr > 0.0 ? a / rr : 1.0;
    sn[j] = rr > 0.0 ? b / rr : 0.0;
    R[j][j] = rr;
    R[j + 1][j] = 0.0;
    const double t = g[j];
    g[j] = cs[j] * t;
    g[j + 1] = -sn[j] * t;
  }
  for (int i = k - 1; i >= 0; --i) {
    double a = g[i];
    for (int c = i + 1; c < k; ++c) a -= R[i][c] * yd[c];
    yd[i] = R[i][i] != 0.0 ? a / R[i][i] : 0.0;
    y[i] = (float)yd[i];
  }
}

constexpr int kInnerMax = 31;

// z = approximately inv(K) v: `its` right-preconditioned FP32 GMRES
// iterations on K32 (block-Jacobi dinv32 inside), from zero, v scaled to
// unit norm for the FP32 range and the result scaled back.
static int inner_solve(IterWork& w, const GmresSetup& g, const double* v, double* z, hipStream_t s) {
  const int n = w.n, N = n / 11, its = std::min(g.inner_iters, kInnerMax);
  const size_t ld = n;
  float* V = w.V32;                          // its+1 vectors
  float* Zs = w.V32 + (size_t)(kInnerMax + 1) * ld;  // its preconditioned vectors
  float* t0 = w.w32;                         // work: 4 n
  float* H = w.h32;                          // (kInnerMax+1) x kInnerMax Hessenberg, then beta, then y
  float* beta = w.h32 + (kInnerMax + 1) * kInnerMax;
  float* y = beta + 1;
  const dim3 blk(256), grd((n + 255) / 256);
  // beta = ||v||, V0 = v / beta (in FP32)
  RSD_TRY(to_f32(v, t0, n, s));
  hipLaunchKernelGGL(multi_dot_kernel<float>, dim3(1), blk, 0, s, (const float*)t0, ld, n, (const float*)t0, beta);
  hipLaunchKernelGGL(sqrt_kernel<float>, dim3(1), dim3(64), 0, s, beta);
  hipLaunchKernelGGL(scale_kernel<float>, grd, blk, 0, s, n, (const float*)t0, V, (const float*)beta, 1.0f, 1);
  for (int j = 0; j < its; ++j) {
    float* Vj = V + (size_t)j * ld;
    float* Zj = Zs + (size_t)j * ld;
    float* wv = V + (size_t)(j + 1) * ld;
    float* hj = H + (size_t)j * (kInnerMax + 1);
    if (g.precond == kIterMixedLearned) RSD_TRY(block_apply<float>(N, w.dinv32, w.corr, g.Lcorr, Vj, Zj, s));
    else RSD_TRY(block_apply<float>(N, w.dinv32, nullptr, nullptr, Vj, Zj, s));
    RSD_TRY(apply_k<float>(*g.L32, g.M32, g.ld32, Zj, wv, t0, w.part32, s));
    // classical Gram-Schmidt, twice
    hipLaunchKernelGGL(multi_dot_kernel<float>, dim3(j + 1), blk, 0, s, (const float*)V, ld, n, (const float*)wv, hj);
    hipLaunchKernelGGL(sub_vh_kernel<float>, grd, blk, 0, s, (const float*)V, ld, n, j + 1, (const float*)hj, wv);
    float* h2 = t0 + 2 * (size_t)n;  // scratch column
    hipLaunchKernelGGL(multi_dot_kernel<float>, dim3(j + 1), blk, 0, s, (const float*)V, ld, n, (const float*)wv, h2);
    hipLaunchKernelGGL(sub_vh_kernel<float>, grd, blk, 0, s, (const float*)V, ld, n, j + 1, (const float*)h2, wv);
    // hj += h2 (j+1 entries), h_{j+1} = ||w||, V_{j+1} = w / h_{j+1}
    hipLaunchKernelGGL(vec_add_kernel<float>, dim3(1), dim3(64), 0, s, j + 1, (const float*)h2, hj);
    hipLaunchKernelGGL(multi_dot_kernel<float>, dim3(1), blk, 0, s, (const float*)wv, ld, n, (const float*)wv,
                       hj + j + 1);
    hipLaunchKernelGGL(sqrt_kernel<float>, dim3(1), dim3(64), 0, s, hj + j + 1);
    hipLaunchKernelGGL(scale_kernel<float>, grd, blk, 0, s, n, (const float*)wv, wv, (const float*)(hj + j + 1), 1.0f,
                       1);
  }
  hipLaunchKernelGGL(inner_lsq_kernel, dim3(1), dim3(64), 0, s, (const float*)H, kInnerMax + 1, its,
                     (const float*)beta, y);
  // z = beta-scaled Z y: t0 = Z y, then z = (double) t0
  hipLaunchKernelGGL(scale_kernel<float>, grd, blk, 0, s, n, (const float*)t0, t0, (const float*)nullptr, 0.0f, 0);
  hipLaunchKernelGGL(add_zy_kernel<float>, grd, blk, 0, s, (const float*)Zs, ld, n, its, (const float*)y, t0);
  RSD_TRY(to_f64(t0, z, n, s));
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

__global__ void copy_kernel(int n, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// inner_solve on the fixed buffers w.vin -> w.zout, captured once into a HIP
// graph and replayed: its ~13 launches an iteration cost the host one graph
// launch.  Everything it reads is fixed by the key (the matrices, the
// preconditioner's blocks, the iteration count), and it records no memset.
static int inner_solve_graph(IterWork& w, const GmresSetup& g, const double* v, double* z, hipStream_t s) {
  const int n = w.n;
  const void* key[8] = {g.M32, g.L32->blk, g.L32->self, g.L32->nnb, w.dinv32, w.corr,
                        g.Lcorr ? (const void*)g.Lcorr->nnb : nullptr, w.V32};
  const int key_i[4] = {n, std::min(g.inner_iters, kInnerMax), g.precond, g.ld32};
  bool same = w.inner_exec != nullptr;
  for (int k = 0; k < 8 && same; ++k) same = key[k] == w.inner_key[k];
  for (int k = 0; k < 4 && same; ++k) same = key_i[k] == w.inner_key_i[k];
  if (!same) {
    if (w.inner_exec) (void)hipGraphExecDestroy(w.inner_exec);
    if (w.inner_graph) (void)hipGraphDestroy(w.inner_graph);
    w.inner_exec = nullptr;
    w.inner_graph = nullptr;
    RSD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = inner_solve(w, g, w.vin, w.zout, s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc != RSD_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    RSD_HIP(e);
    w.inner_graph = graph;
    RSD_HIP(hipGraphInstantiate(&w.inner_exec, graph, nullptr, nullptr, 0));
    for (int k = 0; k < 8; ++k) w.inner_key[k] = key[k];
    for (int k = 0; k < 4; ++k) w.inner_key_i[k] = key_i[k];
  }
  const dim3 blk(256), grd((n + 255) / 256);
  hipLaunchKernelGGL(copy_kernel, grd, blk, 0, s, n, v, w.vin);
  // one replay streams the FP32 M through the gemv once per inner iteration
  const double its = key_i[1], mb = (double)g.ld32 * n * sizeof(float);
  hipEvent_t pe = prof_begin(kProfInnerGraph, its * 2.0 * (double)n * n, its * mb, s);
  RSD_HIP(hipGraphLaunch(w.inner_exec, s));
  prof_end(pe, s);
  hipLaunchKernelGGL(copy_kernel, grd, blk, 0, s, n, (const double*)w.zout, z);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------------------- the FP64 outer FGMRES
int iter_reserve(IterWork& w, int n11, int m, int N, int bsr_cap) {
  if (m > kInnerMax) m = kInnerMax;
  if ((size_t)n11 <= w.cap_n && w.m == m) {
    w.n = n11;
    return RSD_OK;
  }
  iter_release(w);
  const size_t n = n11;
  RSD_HIP(hipMalloc(&w.V, (size_t)(m + 1) * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.Z, (size_t)m * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.w, 4 * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.part, ((n + kGemvCols - 1) / kGemvCols) * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.h, 2 * (size_t)(m + 2) * sizeof(double)));
  RSD_HIP(hipHostMalloc(&w.hh, 2 * (size_t)(m + 2) * sizeof(double)));
  RSD_HIP(hipMalloc(&w.V32, (size_t)(2 * kInnerMax + 1) * n * sizeof(float)));
  RSD_HIP(hipMalloc(&w.w32, 4 * n * sizeof(float)));
  RSD_HIP(hipMalloc(&w.part32, ((n + kGemvCols - 1) / kGemvCols) * n * sizeof(float)));
  RSD_HIP(hipMalloc(&w.h32, ((kInnerMax + 1) * kInnerMax + 2 + kInnerMax) * sizeof(float)));
  RSD_HIP(hipMalloc(&w.dinv, 121 * (size_t)N * sizeof(double)));
  RSD_HIP(hipMalloc(&w.dinv32, 121 * (size_t)N * sizeof(float)));
  RSD_HIP(hipMalloc(&w.corr, 121 * (size_t)N * (bsr_cap + 1) * sizeof(double)));
  RSD_HIP(hipMalloc(&w.vin, n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.zout, n * sizeof(double)));
  w.cap_n = n;
  w.n = n11;
  w.m = m;
  return RSD_OK;
}

void iter_release(IterWork& w) {
  if (w.inner_exec) (void)hipGraphExecDestroy(w.inner_exec);
  if (w.inner_graph) (void)hipGraphDestroy(w.inner_graph);
  for (void* p : {(void*)w.V, (void*)w.Z, (void*)w.w, (void*)w.part, (void*)w.h, (void*)w.V32, (void*)w.w32,
                  (void*)w.part32, (void*)w.h32, (void*)w.dinv, (void*)w.dinv32, (void*)w.corr, (void*)w.vin,
                  (void*)w.zout})
    if (p) (void)hipFree(p);
  if (w.hh) (void)hipHostFree(w.hh);
  w = IterWork{};
}

static int precondition(IterWork& w, const GmresSetup& g, const double* v, double* z, hipStream_t s) {
  const int n = w.n, N = n / 11;
  switch (g.precond) {
    case kIterNone:
      RSD_HIP(hipMemcpyAsync(z, v, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      return RSD_OK;
    case kIterBlockJacobi:
      return block_apply<double>(N, w.dinv, nullptr, nullptr, v, z, s);
    case kIterLearned:
      return block_apply<double>(N, w.dinv, w.corr, g.Lcorr, v, z, s);
    case kIterMixed:
    case kIterMixedLearned:
      return g.graphs ? inner_solve_graph(w, g, v, z, s) : inner_solve(w, g, v, z, s);
  }
  set_error("gmres: unknown preconditioner");
  return RSD_ERR_ARG;
}

int gmres_solve(IterWork& w, const GmresSetup& g, const double* b, double* x, double tol, int max_iter,
                GmresResult* res, hipStream_t s) {
  const int n = w.n, m = w.m;
  const size_t ld = n;
  double* r = w.w;             // residual / work
  double* y = w.w + n;         // K x
  double* kw = w.w + 2 * n;    // apply_k scratch (2 n)
  const dim3 blk(256), grd((n + 255) / 256);
  auto kx = [&](const double* in, double* out) { return apply_k<double>(*g.L, g.M, g.ld, in, out, kw, w.part, s); };
  auto norm = [&](const double* v, double* out) -> int {
    hipLaunchKernelGGL(multi_dot_kernel<double>, dim3(1), blk, 0, s, v, ld, n, v, out);
    RSD_HIP(hipGetLastError());
    return RSD_OK;
  };
  // ||b||
  RSD_TRY(norm(b, w.h));
  RSD_HIP(hipMemcpyAsync(w.hh, w.h, sizeof(double), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(w.hh[0]);
  *res = GmresResult{};
  if (!(bnorm > 0.0)) {
    RSD_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    res->converged = true;
    return RSD_OK;
  }
  int it = 0;
  double prev_beta = 0.0;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gg(m + 1), yy(m);
  while (true) {
    RSD_TRY(kx(x, y));
    hipLaunchKernelGGL(residual_kernel, grd, blk, 0, s, n, b, (const double*)y, r);
    RSD_TRY(norm(r, w.h));
    RSD_HIP(hipMemcpyAsync(w.hh, w.h, sizeof(double), hipMemcpyDeviceToHost, s));
    RSD_HIP(hipStreamSynchronize(s));
    const double beta = std::sqrt(w.hh[0]);
    res->relres = beta / bnorm;
    if (beta <= tol * bnorm) {
      res->converged = true;
      break;
    }
    if (!std::isfinite(beta) || it >= max_iter) break;
    // near the tolerance, a cycle that did not halve the true residual has
    // met FP64's floor for this system (the rounding of R_lub (M x), whose
    // entries grow as 1/gap): within 100 tol that is the answer, else give up
    // rather than spin to max_iter
    if (prev_beta > 0.0 && beta > 0.5 * prev_beta && beta < 1e3 * tol * bnorm) {
      res->converged = beta <= 1e2 * tol * bnorm;
      res->at_floor = res->converged;
      break;
    }
    prev_beta = beta;
    // V0 = r / beta
    double* hb = w.h + (m + 2);
    w.hh[m + 2] = beta;
    RSD_HIP(hipMemcpyAsync(hb, w.hh + (m + 2), sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(scale_kernel<double>, grd, blk, 0, s, n, (const double*)r, w.V, (const double*)hb, 1.0, 1);
    std::fill(gg.begin(), gg.end(), 0.0);
    gg[0] = beta;
    int k = 0;
    bool done = false;
    for (; k < m && it < max_iter; ++k, ++it) {
      double* Vk = w.V + (size_t)k * ld;
      double* Zk = w.Z + (size_t)k * ld;
      double* wv = w.V + (size_t)(k + 1) * ld;
      RSD_TRY(precondition(w, g, Vk, Zk, s));
      RSD_TRY(kx(Zk, wv));
      double* h1 = w.h;
      double* h2 = w.h + (m + 2);
      hipLaunchKernelGGL(multi_dot_kernel<double>, dim3(k + 1), blk, 0, s, (const double*)w.V, ld, n,
                         (const double*)wv, h1);
      hipLaunchKernelGGL(sub_vh_kernel<double>, grd, blk, 0, s, (const double*)w.V, ld, n, k + 1,
                         (const double*)h1, wv);
      hipLaunchKernelGGL(multi_dot_kernel<double>, dim3(k + 1), blk, 0, s, (const double*)w.V, ld, n,
                         (const double*)wv, h2);
      hipLaunchKernelGGL(sub_vh_kernel<double>, grd, blk, 0, s, (const double*)w.V, ld, n, k + 1,
                         (const double*)h2, wv);
      hipLaunchKernelGGL(multi_dot_kernel<double>, dim3(1), blk, 0, s, (const double*)wv, ld, n, (const double*)wv,
                         h2 + k + 1);
      RSD_HIP(hipMemcpyAsync(w.hh, w.h, 2 * (size_t)(m + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
      RSD_HIP(hipStreamSynchronize(s));
      double* hc = &H[(size_t)k * (m + 1)];
      for (int i = 0; i <= k; ++i) hc[i] = w.hh[i] + w.hh[(m + 2) + i];
      const double hn = std::sqrt(w.hh[(m + 2) + k + 1]);
      hc[k + 1] = hn;
      // V_{k+1} = w / hn
      w.hh[(m + 2) + k + 1] = hn;
      RSD_HIP(hipMemcpyAsync(h2 + k + 1, w.hh + (m + 2) + k + 1, sizeof(double), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(scale_kernel<double>, grd, blk, 0, s, n, (const double*)wv, wv, (const double*)(h2 + k + 1),
                         1.0, 1);
      for (int i = 0; i < k; ++i) {
        const double a = hc[i], bb = hc[i + 1];
        hc[i] = cs[i] * a + sn[i] * bb;
        hc[i + 1] = -sn[i] * a + cs[i] * bb;
      }
      const double a = hc[k], bb = hc[k + 1], rr = std::sqrt(a * a + bb * bb);
      cs[k] = rr > 0.0 ? a / rr : 1.0;
      sn[k] = rr > 0.0 ? bb / rr : 0.0;
      hc[k] = rr;
      hc[k + 1] = 0.0;
      gg[k + 1] = -sn[k] * gg[k];
      gg[k] = cs[k] * gg[k];
      if (!std::isfinite(gg[k + 1])) {
        done = true;
        ++k;
        ++it;
        break;
      }
      if (std::fabs(gg[k + 1]) <= 0.5 * tol * bnorm) {  // the estimate; the true residual decides below
        ++k;
        ++it;
        done = true;
        break;
      }
    }
    // x += Z y, y from the triangular system
    for (int i = k - 1; i >= 0; --i) {
      double acc = gg[i];
      for (int c = i + 1; c < k; ++c) acc -= H[(size_t)c * (m + 1) + i] * yy[c];
      yy[i] = H[(size_t)i * (m + 1) + i] != 0.0 ? acc / H[(size_t)i * (m + 1) + i] : 0.0;
    }
    if (k > 0) {
      RSD_HIP(hipMemcpyAsync(w.h, yy.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(add_zy_kernel<double>, grd, blk, 0, s, (const double*)w.Z, ld, n, k, (const double*)w.h, x);
      RSD_HIP(hipGetLastError());
      RSD_HIP(hipStreamSynchronize(s));  // w.h is reused as the next step's scratch
    }
    (void)done;
  }
  res->iterations = it;
  return RSD_OK;
}

template int dense_gemv<double>(const double*, int, int, const double*, double*, double*, hipStream_t);
template int dense_gemv<float>(const float*, int, int, const float*, float*, float*, hipStream_t);
template int apply_k<double>(const LubBsr<double>&, const double*, int, const double*, double*, double*, double*,
                             hipStream_t);
template int bsr_apply<double>(const LubBsr<double>&, const double*, double*, bool, hipStream_t);
template int bsr_apply<float>(const LubBsr<float>&, const float*, float*, bool, hipStream_t);
template int block_apply<double>(int, const double*, const double*, const LubBsr<double>*, const double*, double*,
                                 hipStream_t);
template int block_apply<float>(int, const float*, const double*, const LubBsr<double>*, const float*, float*,
                                hipStream_t);

}  // namespace rsd
