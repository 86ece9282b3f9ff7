// rsd_dense.hip -- FP64 dense linear algebra for the grand-resistance solve.
//
// Replaces MATREV (matrices.f90:4-34: DPOTRF('U') + DPOTRI('U') + mirror)
// with a right-looking blocked Cholesky, a recursive triangular inverse and
// a triangular X^T X product, all built on one batched 64x64-tile GEMM that
// runs on the CDNA4 FP64 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Layout: column-major, order padded to a multiple of 64 with an identity
// block so no kernel needs edge handling.  The symmetric input is read from
// its lower triangle; the reference reads the upper one, and the callers
// build the lower triangle as the mirror of the reference's upper triangle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "rsd_internal.h"
#include "rsd_prof.h"

namespace rsd {

typedef double v4d __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ GEMM
// One workgroup = one 64x64 output tile, 4 waves in a 2x2 arrangement, each
// wave a 32x32 sub-tile = 2x2 MFMA 16x16 blocks.  Inner dimension staged
// through LDS in chunks of 16.
//
// MFMA f64 16x16x4 operand layout (lane l, 0..63):
//   A: A[l%16][l/16]     B: B[l/16][l%16]     D: D[(l/16)+4*v][l%16], v=0..3
constexpr int KC = 16;
constexpr int LDSP = 64 + 2;  // padded LDS row (doubles)

// Up to kInlineDesc descriptors travel in the kernel arguments (no copy on
// the stream); larger batches use a device array.
constexpr int kInlineDesc = 20;
struct GemmArgs {
  GemmDesc d[kInlineDesc];
  const GemmDesc* ext;
  int nd, flags;
};

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm64_kernel(GemmArgs args) {
  const GemmDesc* __restrict__ descs = (args.nd <= kInlineDesc) ? args.d : args.ext;
  const int ndesc = args.nd, flags = args.flags;
  __shared__ double As[KC][LDSP];
  __shared__ double Bs[KC][LDSP];
  __shared__ double Cs[64][LDSP];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int t = blockIdx.x;
  int d = 0;
  while (d + 1 < ndesc && descs[d + 1].tile_base <= t) ++d;
  const GemmDesc D = descs[d];
  const int local = t - D.tile_base;
  int ti, tj;
  if (flags & kLowerC) {
    ti = (int)((sqrt(8.0 * local + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= local) ++ti;
    while (ti * (ti + 1) / 2 > local) --ti;
    tj = local - ti * (ti + 1) / 2;
  } else {
    ti = local % D.mt;
    tj = local / D.mt;
  }
  int klo = 0, khi = D.kt;
  if (flags & kKloRow) klo = max(klo, ti);
  if (flags & kKloCol) klo = max(klo, tj);
  if (flags & kKhiRow) khi = min(khi, ti + 1);

  const int m0 = ti * 64, n0 = tj * 64;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  v4d acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

  const double* __restrict__ A = D.A;
  const double* __restrict__ B = D.B;
  for (int kb = klo * 64; kb < khi * 64; kb += KC) {
    // stage op(A)[m0:m0+64, kb:kb+16] and op(B)[kb:kb+16, n0:n0+64]
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (!TA) {
        int m = tid & 63, k = (tid >> 6) + 4 * r;
        As[k][m] = A[(size_t)(m0 + m) + (size_t)(kb + k) * D.lda];
      } else {
        int k = tid & 15, m = (tid >> 4) + 16 * r;
        As[k][m] = A[(size_t)(kb + k) + (size_t)(m0 + m) * D.lda];
      }
      if (!TB) {
        int k = tid & 15, n = (tid >> 4) + 16 * r;
        Bs[k][n] = B[(size_t)(kb + k) + (size_t)(n0 + n) * D.ldb];
      } else {
        int n = tid & 63, k = (tid >> 6) + 4 * r;
        Bs[k][n] = B[(size_t)(n0 + n) + (size_t)(kb + k) * D.ldb];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a0 = As[kr][wm + (lane & 15)];
      double a1 = As[kr][wm + 16 + (lane & 15)];
      double b0 = Bs[kr][wn + (lane & 15)];
      double b1 = Bs[kr][wn + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // stage the tile through LDS for column-contiguous stores
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r = wm + bm * 16 + (lane >> 4) + 4 * v;
        int c = wn + bn * 16 + (lane & 15);
        Cs[c][r] = acc[bm][bn][v];
      }
  __syncthreads();
  double* __restrict__ Cg = D.C;
  for (int idx = tid; idx < 4096; idx += 256) {
    int r = idx & 63, c = idx >> 6;
    size_t off = (size_t)(m0 + r) + (size_t)(n0 + c) * D.ldc;
    double v = D.alpha * Cs[c][r];
    if (D.beta != 0.0) v += D.beta * Cg[off];
    Cg[off] = v;
  }
}

// inner-dimension tiles actually multiplied, summed over the output tiles
static double desc_ktiles(const GemmDesc& g, int flags) {
  double tot = 0.0;
  const int ntile = (flags & kLowerC) ? g.mt : g.nt;
  for (int ti = 0; ti < g.mt; ++ti)
    for (int tj = 0; tj < ntile; ++tj) {
      if ((flags & kLowerC) && tj > ti) continue;
      int klo = 0, khi = g.kt;
      if (flags & kKloRow) klo = std::max(klo, ti);
      if (flags & kKloCol) klo = std::max(klo, tj);
      if (flags & kKhiRow) khi = std::min(khi, ti + 1);
      if (khi > klo) tot += khi - klo;
    }
  return tot;
}

static int desc_tiles(const GemmDesc& g, int flags) {
  if (flags & kLowerC) return g.mt * (g.mt + 1) / 2;
  return g.mt * g.nt;
}

int gemm_batched(DenseWork& w, const std::vector<GemmDesc>& in, bool ta, bool tb, int flags, hipStream_t s) {
  if (in.empty()) return RSD_OK;
  std::vector<GemmDesc> descs;
  descs.reserve(in.size());
  int total = 0;
  for (GemmDesc g : in) {
    if (g.mt <= 0 || g.nt <= 0) continue;
    g.tile_base = total;
    total += desc_tiles(g, flags);
    descs.push_back(g);
  }
  if (descs.empty()) return RSD_OK;
  GemmArgs args{};
  args.nd = (int)descs.size();
  args.flags = flags;
  if (descs.size() <= (size_t)kInlineDesc) {
    for (size_t i = 0; i < descs.size(); ++i) args.d[i] = descs[i];
  } else {
    if (descs.size() > w.desc_cap) {
      set_error("gemm_batched: descriptor buffer too small");
      return RSD_ERR_CAPACITY;
    }
    RSD_HIP(hipMemcpyAsync(w.d_desc, descs.data(), descs.size() * sizeof(GemmDesc), hipMemcpyHostToDevice, s));
    args.ext = w.d_desc;
  }
  dim3 grid(total), block(256);
  hipEvent_t pe = nullptr;
  if (prof_on()) {
    double kt = 0.0;
    for (const GemmDesc& g : descs) kt += desc_ktiles(g, flags);
    pe = prof_begin(kProfGemm, kt * 2.0 * 64.0 * 64.0 * 64.0, 0.0, s);
  }
  if (!ta && !tb)
    hipLaunchKernelGGL((gemm64_kernel<false, false>), grid, block, 0, s, args);
  else if (!ta && tb)
    hipLaunchKernelGGL((gemm64_kernel<false, true>), grid, block, 0, s, args);
  else if (ta && !tb)
    hipLaunchKernelGGL((gemm64_kernel<true, false>), grid, block, 0, s, args);
  else
    hipLaunchKernelGGL((gemm64_kernel<true, true>), grid, block, 0, s, args);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------------ 64x64 diagonal factor
// One workgroup (4 waves) per 64x64 diagonal block.  Thread (w, l) keeps
// row l, columns 16w..16w+15 in registers.  At pivot j the wave owning
// column j scales it and publishes it through a double-buffered LDS column;
// one barrier per pivot.  The triangular inverse (into dinv, column-major)
// is then built blockwise from 16x16 pieces: X_ii = inv(L_ii) by one wave
// each, X_ij = -X_ii sum_k L_ik X_kj level by level.
__global__ __launch_bounds__(256) void potrf_diag64_kernel(double* __restrict__ A, int ld, int k0,
                                                           double* __restrict__ dinv, int* __restrict__ info) {
  __shared__ double col[2][64];
  __shared__ double Ls[64][65];
  __shared__ double Xs[64][65];
  __shared__ double Ts[4][16][17];
  __shared__ int bad_s;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  if (tid == 0) bad_s = 0x7fffffff;
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = A[(size_t)(k0 + l) + (size_t)(k0 + 16 * w + q) * ld];
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    const int wj = j >> 4, qj = j & 15;
    double* cb = col[j & 1];
    if (w == wj) {
      double dj = __shfl(a[qj], j, 64);
      if (!(dj > 0.0) && l == 0) atomicMin(&bad_s, k0 + j + 1);
      const double piv = sqrt(dj), inv = 1.0 / piv;
      const double v = (l > j) ? a[qj] * inv : (l == j ? piv : 0.0);
      a[qj] = v;
      cb[l] = v;
    }
    __syncthreads();
    if (w >= wj) {
      const double cl = cb[l];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (16 * w + q > j) a[q] -= cl * cb[16 * w + q];
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 16 * w + q;
    Ls[l][c] = (l >= c) ? a[q] : 0.0;
    Xs[l][c] = 0.0;
  }
  __syncthreads();
  // diagonal 16x16 inverses: wave b, lane c (< 16) computes column c of X_bb
  {
    const int b = w, c = l;
    if (c < 16) {
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) acc += Ls[16 * b + i][16 * b + k] * x[k];
        const double lii = Ls[16 * b + i][16 * b + i];
        x[i] = (i == c) ? 1.0 / lii : ((i > c) ? -acc / lii : 0.0);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) Xs[16 * b + i][16 * b + c] = x[i];
    }
  }
  __syncthreads();
  // off-diagonal blocks at block distance d: X_ij = -X_ii (sum_{k=j}^{i-1} L_ik X_kj)
  for (int d = 1; d < 4; ++d) {
    const int nb = 4 - d;
    if (w < nb) {
      const int j = w, i = w + d;
#pragma unroll
      for (int e = 0; e < 4; ++e) {  // 256 entries of T per wave, 4 per lane
        const int idx = l + 64 * e, r = idx & 15, c = idx >> 4;
        double acc = 0.0;
        for (int k = 16 * j; k < 16 * i; ++k) acc += Ls[16 * i + r][k] * Xs[k][16 * j + c];
        Ts[w][r][c] = acc;
      }
    }
    __syncthreads();
    if (w < nb) {
      const int j = w, i = w + d;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = l + 64 * e, r = idx & 15, c = idx >> 4;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += Xs[16 * i + r][16 * i + k] * Ts[w][k][c];
        Xs[16 * i + r][16 * j + c] = -acc;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 16 * w + q;
    A[(size_t)(k0 + l) + (size_t)(k0 + c) * ld] = Ls[l][c];
    dinv[(size_t)l + (size_t)c * 64] = Xs[l][c];
  }
  if (tid == 0 && bad_s != 0x7fffffff) atomicMin(info, bad_s);
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

int potrf_lower(DenseWork& w, double* A, int n, int ld, hipStream_t s) {
  const int nt = n / 64;
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, w.d_info, 0x7fffffff);
  for (int k = 0; k < nt; ++k) {
    const int k0 = 64 * k;
    hipEvent_t pe = prof_begin(kProfPotrfDiag, 2.0 * 64.0 * 64.0 * 64.0 / 3.0, 0.0, s);
    hipLaunchKernelGGL(potrf_diag64_kernel, dim3(1), dim3(256), 0, s, A, ld, k0, w.dinv + (size_t)k * 4096,
                       w.d_info);
    prof_end(pe, s);
    const int below = nt - k - 1;
    if (below > 0) {
      // panel solve A_ik <- A_ik Dinv_k^T on the matrix cores (in place: a
      // workgroup owns all 64 columns of its row tile and reads them before writing)
      GemmDesc t{};
      t.A = A + (size_t)(k0 + 64) + (size_t)k0 * ld;
      t.B = w.dinv + (size_t)k * 4096;
      t.C = A + (size_t)(k0 + 64) + (size_t)k0 * ld;
      t.lda = ld;
      t.ldb = 64;
      t.ldc = ld;
      t.mt = below;
      t.nt = 1;
      t.kt = 1;
      t.alpha = 1.0;
      t.beta = 0.0;
      int rc = gemm_batched(w, {t}, false, true, 0, s);
      if (rc) return rc;
      // trailing update A22 -= P P^T (lower tiles)
      GemmDesc g{};
      g.A = A + (size_t)(k0 + 64) + (size_t)k0 * ld;
      g.B = g.A;
      g.C = A + (size_t)(k0 + 64) + (size_t)(k0 + 64) * ld;
      g.lda = g.ldb = g.ldc = ld;
      g.mt = g.nt = below;
      g.kt = 1;
      g.alpha = -1.0;
      g.beta = 1.0;
      rc = gemm_batched(w, {g}, false, true, kLowerC, s);
      if (rc) return rc;
    }
  }
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// copy the stored diagonal inverses onto the diagonal tiles
__global__ __launch_bounds__(256) void put_dinv_kernel(double* A, int ld, const double* dinv) {
  const int k = blockIdx.x;
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    int r = idx & 63, c = idx >> 6;
    A[(size_t)(64 * k + r) + (size_t)(64 * k + c) * ld] = dinv[(size_t)k * 4096 + r + (size_t)c * 64];
  }
}

// Recursive blocked inverse of a lower-triangular matrix:
//   inv([L11 0; L21 L22]) = [X11 0; -X22 L21 X11  X22]
// The leaves are the 64x64 inverses potrf_lower left in dinv; every level
// of the recursion is one batched launch for T = L21 X11 and one for
// X21 = -X22 T.
namespace {
struct Node {
  int a, m, b;  // tile range [a,b) split at m
};
void build_levels(int a, int b, int depth, std::vector<std::vector<Node>>& lv) {
  if (b - a <= 1) return;
  int m = a + (b - a + 1) / 2;
  if ((int)lv.size() <= depth) lv.resize(depth + 1);
  lv[depth].push_back({a, m, b});
  build_levels(a, m, depth + 1, lv);
  build_levels(m, b, depth + 1, lv);
}
}  // namespace

int trtri_lower(DenseWork& w, double* A, int n, int ld, hipStream_t s) {
  const int nt = n / 64;
  hipLaunchKernelGGL(put_dinv_kernel, dim3(nt), dim3(256), 0, s, A, ld, (const double*)w.dinv);
  std::vector<std::vector<Node>> lv;
  build_levels(0, nt, 0, lv);
  double* T = w.tmp;  // same shape as A (leading dim n); node blocks never overlap
  for (int d = (int)lv.size() - 1; d >= 0; --d) {
    std::vector<GemmDesc> g1, g2;
    for (const Node& nd : lv[d]) {
      GemmDesc g{};
      // T[m:b, a:m] = L21 * X11   (X11 lower: inner index starts at the column tile)
      g.A = A + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * ld;
      g.B = A + (size_t)(64 * nd.a) + (size_t)(64 * nd.a) * ld;
      g.C = T + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * n;
      g.lda = ld;
      g.ldb = ld;
      g.ldc = n;
      g.mt = nd.b - nd.m;
      g.nt = nd.m - nd.a;
      g.kt = nd.m - nd.a;
      g.alpha = 1.0;
      g.beta = 0.0;
      g1.push_back(g);
      GemmDesc h{};
      // X21 = -X22 * T   (X22 lower: inner index ends at the row tile)
      h.A = A + (size_t)(64 * nd.m) + (size_t)(64 * nd.m) * ld;
      h.B = T + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * n;
      h.C = A + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * ld;
      h.lda = ld;
      h.ldb = n;
      h.ldc = ld;
      h.mt = nd.b - nd.m;
      h.nt = nd.m - nd.a;
      h.kt = nd.b - nd.m;
      h.alpha = -1.0;
      h.beta = 0.0;
      g2.push_back(h);
    }
    int rc = gemm_batched(w, g1, false, false, kKloCol, s);
    if (rc) return rc;
    rc = gemm_batched(w, g2, false, false, kKhiRow, s);
    if (rc) return rc;
  }
  return RSD_OK;
}

int lauum_lower(DenseWork& w, const double* X, int ldx, double* C, int ldc, int n, hipStream_t s) {
  const int nt = n / 64;
  GemmDesc g{};
  g.A = X;
  g.B = X;
  g.C = C;
  g.lda = ldx;
  g.ldb = ldx;
  g.ldc = ldc;
  g.mt = g.nt = g.kt = nt;
  g.alpha = 1.0;
  g.beta = 0.0;
  // C_ij = sum_{k >= i} X_ki^T X_kj   (i >= j)
  return gemm_batched(w, {g}, true, false, kLowerC | kKloRow, s);
}

__global__ void symmetrize_kernel(double* A, int n, int ld) {
  const int c = blockIdx.y * 16 + threadIdx.y;
  const int r = blockIdx.x * 16 + threadIdx.x;
  if (r < n && c < n && r < c) A[(size_t)r + (size_t)c * ld] = A[(size_t)c + (size_t)r * ld];
}

int symmetrize_from_lower(double* A, int n, int ld, hipStream_t s) {
  dim3 b(16, 16), g((n + 15) / 16, (n + 15) / 16);
  hipLaunchKernelGGL(symmetrize_kernel, g, b, 0, s, A, n, ld);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

__global__ void load_sym_upper_kernel(const double* src, int n, int lds, double* A, int np, int ld) {
  const int c = blockIdx.y * 16 + threadIdx.y;
  const int r = blockIdx.x * 16 + threadIdx.x;
  if (r >= np || c >= np) return;
  double v;
  if (r < n && c < n)
    v = (r <= c) ? src[(size_t)r + (size_t)c * lds] : src[(size_t)c + (size_t)r * lds];
  else
    v = (r == c) ? 1.0 : 0.0;
  A[(size_t)r + (size_t)c * ld] = v;
}

int load_sym_from_upper(const double* src, int n, int ldsrc, double* A, int np, int ld, hipStream_t s) {
  dim3 b(16, 16), g((np + 15) / 16, (np + 15) / 16);
  hipLaunchKernelGGL(load_sym_upper_kernel, g, b, 0, s, src, n, ldsrc, A, np, ld);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

__global__ void identity_pad_kernel(double* A, int n, int np, int ld) {
  const int c = blockIdx.y * 16 + threadIdx.y;
  const int r = blockIdx.x * 16 + threadIdx.x;
  if (r >= np || c >= np) return;
  if (r < n && c < n) return;
  A[(size_t)r + (size_t)c * ld] = (r == c) ? 1.0 : 0.0;
}

int set_identity_padding(double* A, int n, int np, int ld, hipStream_t s) {
  if (np == n) return RSD_OK;
  dim3 b(16, 16), g((np + 15) / 16, (np + 15) / 16);
  hipLaunchKernelGGL(identity_pad_kernel, g, b, 0, s, A, n, np, ld);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

__global__ void copy_block_kernel(const double* src, int lds, double* dst, int ldd, int m, int n) {
  const int c = blockIdx.y;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x)
    if (c < n) dst[(size_t)r + (size_t)c * ldd] = src[(size_t)r + (size_t)c * lds];
}

int copy_block(const double* src, int lds, double* dst, int ldd, int m, int n, hipStream_t s) {
  if (m <= 0 || n <= 0) return RSD_OK;
  dim3 g(std::max(1, std::min(64, (m + 255) / 256)), n), b(256);
  hipLaunchKernelGGL(copy_block_kernel, g, b, 0, s, src, lds, dst, ldd, m, n);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

__global__ void matrix_axpy_kernel(double alpha, const double* X, int ldx, double* Y, int ldy, int m, int n) {
  const int c = blockIdx.y;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x)
    if (c < n) Y[(size_t)r + (size_t)c * ldy] += alpha * X[(size_t)r + (size_t)c * ldx];
}

int matrix_axpy(double alpha, const double* X, int ldx, double* Y, int ldy, int m, int n, hipStream_t s) {
  if (m <= 0 || n <= 0) return RSD_OK;
  dim3 g(std::max(1, std::min(64, (m + 255) / 256)), n), b(256);
  hipLaunchKernelGGL(matrix_axpy_kernel, g, b, 0, s, alpha, X, ldx, Y, ldy, m, n);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int spd_inverse(DenseWork& w, double* A, double* Ainv, int np, int ld, hipStream_t s) {
  int rc = potrf_lower(w, A, np, ld, s);
  if (rc) return rc;
  rc = trtri_lower(w, A, np, ld, s);
  if (rc) return rc;
  rc = lauum_lower(w, A, ld, Ainv, ld, np, s);
  if (rc) return rc;
  return symmetrize_from_lower(Ainv, np, ld, s);
}

// ------------------------------------------------------------------ GEMV
// y = alpha*A x + beta*y: 64-row x column-chunk workgroups write partial sums
// (4 waves stride the chunk's columns, lanes run down contiguous rows), then
// a second pass adds the chunks in a fixed order (deterministic).
__global__ __launch_bounds__(256) void gemv_n_part_kernel(const double* __restrict__ A, int lda, int m, int n,
                                                          const double* __restrict__ x, int cchunk,
                                                          double* __restrict__ part) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
  const int c0 = blockIdx.y * cchunk, c1 = min(n, c0 + cchunk);
  double acc = 0.0;
  if (r < m)
    for (int c = c0 + wv; c < c1; c += 4) acc += A[(size_t)r + (size_t)c * lda] * x[c];
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && r < m) part[(size_t)blockIdx.y * m + r] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

__global__ __launch_bounds__(256) void gemv_n_reduce_kernel(int m, int nchunk, const double* __restrict__ part,
                                                            double alpha, double beta, double* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += part[(size_t)c * m + r];
  y[r] = alpha * acc + (beta != 0.0 ? beta * y[r] : 0.0);
}

// y = alpha*A^T x + beta*y : one wavefront per output (column), lanes down the column
__global__ __launch_bounds__(256) void gemv_t_kernel(const double* __restrict__ A, int lda, int m, int n, double alpha,
                                                     const double* __restrict__ x, double beta, double* y) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  double acc = 0.0;
  for (int r = lane; r < m; r += 64) acc += A[(size_t)r + (size_t)c * lda] * x[r];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[c] = alpha * acc + (beta != 0.0 ? beta * y[c] : 0.0);
}

int gemv(DenseWork& w, const double* A, int lda, int m, int n, bool trans, double alpha, const double* x, double beta,
         double* y, hipStream_t s) {
  if (m <= 0 || n <= 0) return RSD_OK;
  hipEvent_t pe = prof_begin(kProfGemv, 2.0 * m * n, 8.0 * m * n, s);
  if (!trans) {
    const int nchunk = std::max(1, std::min(64, (n + 63) / 64));
    const int cchunk = (n + nchunk - 1) / nchunk;
    const size_t need = (size_t)m * nchunk;
    if (need > w.gv_cap) {
      if (w.gv_part) RSD_HIP(hipFree(w.gv_part));
      w.gv_part = nullptr;
      RSD_HIP(hipMalloc(&w.gv_part, need * sizeof(double)));
      w.gv_cap = need;
    }
    hipLaunchKernelGGL(gemv_n_part_kernel, dim3((m + 63) / 64, nchunk), dim3(256), 0, s, A, lda, m, n, x, cchunk,
                       w.gv_part);
    hipLaunchKernelGGL(gemv_n_reduce_kernel, dim3((m + 255) / 256), dim3(256), 0, s, m, nchunk,
                       (const double*)w.gv_part, alpha, beta, y);
  } else {
    hipLaunchKernelGGL(gemv_t_kernel, dim3((n + 3) / 4), dim3(256), 0, s, A, lda, m, n, alpha, x, beta, y);
  }
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------------------------------ scratch
int dense_reserve(DenseWork& w, int n, hipStream_t s) {
  (void)s;
  size_t nt = (size_t)(n / 64);
  size_t need_dinv = std::max<size_t>(nt, 1) * 4096;
  if (need_dinv > w.dinv_cap) {
    if (w.dinv) RSD_HIP(hipFree(w.dinv));
    RSD_HIP(hipMalloc(&w.dinv, need_dinv * sizeof(double)));
    w.dinv_cap = need_dinv;
  }
  size_t need_tmp = (size_t)n * (size_t)n;
  if (need_tmp > w.tmp_cap) {
    if (w.tmp) RSD_HIP(hipFree(w.tmp));
    RSD_HIP(hipMalloc(&w.tmp, need_tmp * sizeof(double)));
    w.tmp_cap = need_tmp;
  }
  if (!w.d_desc) {
    w.desc_cap = 64 * 1024;
    RSD_HIP(hipMalloc(&w.d_desc, w.desc_cap * sizeof(GemmDesc)));
  }
  if (!w.d_info) RSD_HIP(hipMalloc(&w.d_info, sizeof(int)));
  return RSD_OK;
}

void dense_release(DenseWork& w) {
  if (w.dinv) (void)hipFree(w.dinv);
  if (w.tmp) (void)hipFree(w.tmp);
  if (w.d_desc) (void)hipFree(w.d_desc);
  if (w.d_info) (void)hipFree(w.d_info);
  if (w.gv_part) (void)hipFree(w.gv_part);
  w = DenseWork{};
}

}  // namespace rsd
