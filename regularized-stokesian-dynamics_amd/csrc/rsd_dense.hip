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
// wave a 32x32 sub-tile = 2x2 MFMA 16x16 blocks.  The inner dimension goes
// through LDS in chunks of 32, NS = 2 or 3 chunks in flight: a chunk is
// staged by direct 16-byte global->LDS loads (8 per wave), issued NS-1 chunks
// ahead of the products that read it.  NS = 3 (96 KB of LDS, one workgroup
// per CU) hides more latency; NS = 2 (64 KB, two per CU) wins once a launch
// has tiles enough for two rounds of workgroups.
//
// MFMA f64 16x16x4 operand layout (lane l, 0..63):
//   A: A[l%16][l/16]     B: B[l/16][l%16]     D: D[(l/16)+4*v][l%16], v=0..3
constexpr int LDSP = 64 + 2;  // padded LDS row (doubles)
constexpr int GKC = 32;              // inner dimension per staged chunk
constexpr int GCH = 64 * GKC;        // doubles per operand chunk
constexpr int GLD = 8;               // direct loads per wave per chunk (both operands)
constexpr int kWaitVm0 = 0x0f70;     // s_waitcnt vmcnt(0): expcnt and lgkmcnt left at their maxima
constexpr int kCpolSc1 = 16;         // cache policy of a direct-to-LDS load: sc1 (agent scope, coherent)
// s_waitcnt vmcnt(n), n < 64, expcnt and lgkmcnt left at their maxima
constexpr int wait_vm(int n) { return (n & 15) | ((n >> 4) << 14) | 0x0f70; }

// Up to kInlineDesc descriptors travel in the kernel arguments (no copy on
// the stream); larger batches use a device array.
constexpr int kInlineDesc = 20;
struct GemmArgs {
  GemmDesc d[kInlineDesc];
  const GemmDesc* ext;
  int nd, flags;
};

// Staged operand chunks, two layouts (r = the output row or column index,
// k = the inner index within the chunk, 0..31):
//  r-contiguous (op(X)(r, k) = X[r + k ld]: A, or B transposed): element
//    (r, k) at k*64 + (r ^ 16 (k & 1));
//  k-contiguous (op(X)(r, k) = X[k + r ld]: A transposed, or B): element
//    (r, k) at r*32 + 2 ((k/2) ^ (r & 15)) + (k & 1).
// The swaps put the 16 rows (or the 4 inner indices) that one MFMA operand
// read touches on different LDS banks.
template <bool KCONT>
__device__ __forceinline__ int gemm_staged(int r, int k) {
  return KCONT ? r * 32 + 2 * ((k >> 1) ^ (r & 15)) + (k & 1) : k * 64 + (r ^ ((k & 1) << 4));
}

// 64 x 32 chunk of op(X) at (r0, k0), 4 direct loads per wave
template <bool KCONT>
__device__ __forceinline__ void gemm_stage(const double* __restrict__ X, size_t ld, int r0, int k0,
                                           double* __restrict__ dst) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = wave * 4 + u;  // 16 loads of 1 KB: 2 columns (r-contiguous) or 4 rows (k-contiguous)
    const double* src;
    if (KCONT) {
      const int r = 4 * q + (lane >> 4), p = (lane & 15) ^ (r & 15);
      src = X + (size_t)(k0 + 2 * p) + (size_t)(r0 + r) * ld;
    } else {
      const int k = 2 * q + (lane >> 5), p = (lane & 31) ^ ((k & 1) << 3);
      src = X + (size_t)(r0 + 2 * p) + (size_t)(k0 + k) * ld;
    }
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)(dst + q * 128), 16, 0, 0);
  }
}

// one chunk's products (cur), the loads of a later chunk (into nxt, when
// pre) in flight meanwhile; cur and nxt are distinct restrict pointers so
// that the compiler lets the reads of cur pass the loads into nxt
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_chunk(v4d (&acc)[2][2], const GemmDesc& D, int m0, int n0, int kload, bool pre,
                                           const double* __restrict__ cur, double* __restrict__ nxt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  if (pre) {
    gemm_stage<TA>(D.A, (size_t)D.lda, m0, kload, nxt);
    gemm_stage<!TB>(D.B, (size_t)D.ldb, n0, kload, nxt + GCH);
  }
  const double* cb = cur + GCH;
#pragma unroll
  for (int kk = 0; kk < GKC; kk += 4) {
    const int kr = kk + (lane >> 4), m = lane & 15;
    const double a0 = cur[gemm_staged<TA>(wm + m, kr)];
    const double a1 = cur[gemm_staged<TA>(wm + 16 + m, kr)];
    const double b0 = cb[gemm_staged<!TB>(wn + m, kr)];
    const double b1 = cb[gemm_staged<!TB>(wn + 16 + m, kr)];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// LDS reads done and every wave here: a barrier without the fence of
// __syncthreads, which would also wait for the chunks still in flight
__device__ __forceinline__ void gemm_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <bool TA, bool TB, int NS>
__global__ __launch_bounds__(256) void gemm64_kernel(GemmArgs args) {
  static_assert(NS == 2 || NS == 3, "two or three chunks in flight");
  const GemmDesc* __restrict__ descs = (args.nd <= kInlineDesc) ? args.d : args.ext;
  const int ndesc = args.nd, flags = args.flags;
  // NS staged chunks of op(A) and op(B); the output tile reuses the space
  __shared__ __attribute__((aligned(16))) double smem[NS * 2 * GCH];
  double(*Cs)[LDSP] = reinterpret_cast<double(*)[LDSP]>(smem);

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
  } else if (flags & kKhiRow) {
    // longest inner ranges first (they end after the row): rows from the bottom
    ti = D.mt - 1 - local / D.nt;
    tj = local % D.nt;
  } else {
    ti = local % D.mt;
    tj = local / D.mt;
  }
  // the tile is the workgroup's: keep its indices (and the chunk loop) scalar
  ti = __builtin_amdgcn_readfirstlane(ti);
  tj = __builtin_amdgcn_readfirstlane(tj);
  int klo = 0, khi = D.kt;
  if (flags & kKloRow) klo = max(klo, ti);
  if (flags & kKloCol) klo = max(klo, tj);
  if (flags & kKhiRow) khi = min(khi, ti + 1);

  const int m0 = ti * 64, n0 = tj * 64;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  v4d acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

  const int k_begin = __builtin_amdgcn_readfirstlane(klo * 64);
  const int nch = __builtin_amdgcn_readfirstlane(khi > klo ? (khi - klo) * (64 / GKC) : 0);
  if (nch > 0) {
#pragma unroll
    for (int c = 0; c < NS - 1; ++c)
      if (c < nch) {
        gemm_stage<TA>(D.A, (size_t)D.lda, m0, k_begin + c * GKC, smem + c * 2 * GCH);
        gemm_stage<!TB>(D.B, (size_t)D.ldb, n0, k_begin + c * GKC, smem + (c * 2 + 1) * GCH);
      }
    for (int c = 0; c < nch; ++c) {
      // chunk c is in (each wave waits for its own loads; the barrier for
      // everyone's), and chunk c-1's buffer is free for chunk c+NS-1
      const int later = min(NS - 2, nch - 1 - c);  // chunks issued after c
      if (NS == 3 && later == 1)
        __builtin_amdgcn_s_waitcnt(wait_vm(GLD));
      else
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
      gemm_barrier();
      const int cn = c + NS - 1;
      gemm_chunk<TA, TB>(acc, D, m0, n0, k_begin + cn * GKC, cn < nch, smem + (c % NS) * 2 * GCH,
                         smem + (cn % NS) * 2 * GCH);
    }
    gemm_barrier();
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
  const bool two = total >= 1024;  // two rounds of workgroups at two per CU
  void (*kern)(GemmArgs);
  if (!ta && !tb)
    kern = two ? gemm64_kernel<false, false, 2> : gemm64_kernel<false, false, 3>;
  else if (!ta && tb)
    kern = two ? gemm64_kernel<false, true, 2> : gemm64_kernel<false, true, 3>;
  else if (ta && !tb)
    kern = two ? gemm64_kernel<true, false, 2> : gemm64_kernel<true, false, 3>;
  else
    kern = two ? gemm64_kernel<true, true, 2> : gemm64_kernel<true, true, 3>;
  hipLaunchKernelGGL(kern, grid, block, 0, s, args);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// A 64x64 column-major tile (leading dimension ld) into LDS by direct
// 16-byte global->LDS loads, 1 KB (two columns) per wave instruction, 8 per
// wave.  Column c of the staged tile holds row m at c*64 + (m ^ 16 (c & 1)):
// the swap of 16-row halves on odd columns puts the two columns an MFMA
// operand read touches on different LDS banks.
// The thread index through an empty asm: address arithmetic derived from it
// is redone at each call instead of being hoisted out of the persistent
// Cholesky's ticket loop and held in registers (spilled) across the tasks.
__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  __asm__ volatile("" : "+v"(t));
  return t;
}

// Hand-offs between workgroups of one launch (the dataflow Cholesky and the
// triangular solves) write through: the producer stores
// the data with agent-scope relaxed atomic stores, which write through the
// L2 to the device's coherence point; each wave waits for its own stores to
// complete (wt_stores_done) before the workgroup barrier that precedes the
// flag store (wt_publish, relaxed).  The consumer polls the flag and then
// reads the data with agent-scope loads, which do not return stale L2 or L1
// copies.  Neither side needs an agent-scope release or acquire fence, which
// writes back or invalidates the whole XCD's L2: under
// load they made each hand-off 1-2 us slower and made the other tiles on
// the XCD re-read their operands.
__device__ __forceinline__ void wt_store(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double wt_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wt_stores_done() { __builtin_amdgcn_s_waitcnt(kWaitVm0); }
__device__ __forceinline__ void wt_publish(int* flag, int gen) {
  __hip_atomic_store(flag, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// after a flag was seen: keeps the compiler from moving the data loads above
// the poll (the hardware issues them after the polled value is in)
__device__ __forceinline__ void wt_seen() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); }

// A 64x64 tile of the factorization (published by another workgroup) into
// LDS by direct 16-byte loads at agent scope, past any stale cached copy.
__device__ __forceinline__ void stage_tile(const double* __restrict__ src, size_t ld, double* __restrict__ dst) {
  const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int chunk = wave * 8 + q;
    const int c = 2 * chunk + (lane >> 5), pr = (lane & 31) ^ ((c & 1) << 3);
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + (size_t)c * ld + 2 * pr),
                                     (__attribute__((address_space(3))) void*)(dst + chunk * 128), 16, 0, kCpolSc1);
  }
}

// element (m, c) of a staged tile
__device__ __forceinline__ int staged(int m, int c) { return c * 64 + (m ^ ((c & 1) << 4)); }

// ------------------------------------------------------------- FP32 GEMM
// The CG preconditioner's products (triangular inverse and Gram product of an
// FP32 copy of the factor): gemm64_kernel's organisation with
// v_mfma_f32_16x16x4f32 (twice the FP64 MFMA rate) and 4-byte elements.
// Operand lanes as the FP64 instruction; output D[4*(l/16)+v][l%16].
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int FCH = 64 * GKC;  // floats per operand chunk

// Staged FP32 chunks (r = output row or column index, k = inner index within
// the chunk, 0..31), 16-byte units swapped so that the 64 lanes of an MFMA
// operand read touch 64 distinct banks:
//  r-contiguous: element (r, k) at k*64 + (r ^ 16 (k & 3));
//  k-contiguous: element (r, k) at r*32 + 4 ((k/4) ^ ((r/2) & 7)) + (k & 3).
template <bool KCONT>
__device__ __forceinline__ int gemm32_staged(int r, int k) {
  return KCONT ? r * 32 + 4 * ((k >> 2) ^ ((r >> 1) & 7)) + (k & 3) : k * 64 + (r ^ ((k & 3) << 4));
}

// 64 x 32 chunk of op(X) at (r0, k0), 2 direct 16-byte loads per wave
template <bool KCONT>
__device__ __forceinline__ void gemm32_stage(const float* __restrict__ X, size_t ld, int r0, int k0,
                                             float* __restrict__ dst) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = wave * 2 + u;  // 8 loads of 1 KB: 4 columns (r-contiguous) or 8 rows (k-contiguous)
    const float* src;
    if (KCONT) {
      const int r = 8 * q + (lane >> 3), p = (lane & 7) ^ ((r >> 1) & 7);
      src = X + (size_t)(k0 + 4 * p) + (size_t)(r0 + r) * ld;
    } else {
      const int k = 4 * q + (lane >> 4), p = (lane & 15) ^ ((k & 3) << 2);
      src = X + (size_t)(r0 + 4 * p) + (size_t)(k0 + k) * ld;
    }
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)(dst + q * 256), 16, 0, 0);
  }
}

struct Gemm32Desc {
  const float* A;
  const float* B;
  float* C;
  int lda, ldb, ldc;
  int mt, nt, kt;  // C is (mt*64) x (nt*64), inner dimension kt*64
  float alpha, beta;
  int tile_base;
};
struct Gemm32Args {
  Gemm32Desc d[kInlineDesc];
  const Gemm32Desc* ext;
  int nd, flags;
};

template <bool TA, bool TB>
__device__ __forceinline__ void gemm32_chunk(v4f (&acc)[2][2], const Gemm32Desc& D, int m0, int n0, int kload,
                                             bool pre, const float* __restrict__ cur, float* __restrict__ nxt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  if (pre) {
    gemm32_stage<TA>(D.A, (size_t)D.lda, m0, kload, nxt);
    gemm32_stage<!TB>(D.B, (size_t)D.ldb, n0, kload, nxt + FCH);
  }
  const float* cb = cur + FCH;
#pragma unroll
  for (int kk = 0; kk < GKC; kk += 4) {
    const int kr = kk + (lane >> 4), m = lane & 15;
    const float a0 = cur[gemm32_staged<TA>(wm + m, kr)];
    const float a1 = cur[gemm32_staged<TA>(wm + 16 + m, kr)];
    const float b0 = cb[gemm32_staged<!TB>(wn + m, kr)];
    const float b1 = cb[gemm32_staged<!TB>(wn + 16 + m, kr)];
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// three chunks in flight (48 KB of LDS)
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm32_kernel(Gemm32Args args) {
  const Gemm32Desc* __restrict__ descs = (args.nd <= kInlineDesc) ? args.d : args.ext;
  const int ndesc = args.nd, flags = args.flags;
  __shared__ __attribute__((aligned(16))) float smem[3 * 2 * FCH];
  float(*Cs)[65] = reinterpret_cast<float(*)[65]>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int t = blockIdx.x;
  int d = 0;
  while (d + 1 < ndesc && descs[d + 1].tile_base <= t) ++d;
  const Gemm32Desc D = descs[d];
  const int local = t - D.tile_base;
  int ti, tj;
  if (flags & kLowerC) {
    ti = (int)((sqrt(8.0 * local + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= local) ++ti;
    while (ti * (ti + 1) / 2 > local) --ti;
    tj = local - ti * (ti + 1) / 2;
  } else if (flags & kKhiRow) {
    ti = D.mt - 1 - local / D.nt;  // longest inner ranges first
    tj = local % D.nt;
  } else {
    ti = local % D.mt;
    tj = local / D.mt;
  }
  ti = __builtin_amdgcn_readfirstlane(ti);
  tj = __builtin_amdgcn_readfirstlane(tj);
  int klo = 0, khi = D.kt;
  if (flags & kKloRow) klo = max(klo, ti);
  if (flags & kKloCol) klo = max(klo, tj);
  if (flags & kKhiRow) khi = min(khi, ti + 1);
  const int m0 = ti * 64, n0 = tj * 64;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  v4f acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
  const int k_begin = __builtin_amdgcn_readfirstlane(klo * 64);
  const int nch = __builtin_amdgcn_readfirstlane(khi > klo ? (khi - klo) * (64 / GKC) : 0);
  if (nch > 0) {
    for (int c = 0; c < 2 && c < nch; ++c) {
      gemm32_stage<TA>(D.A, (size_t)D.lda, m0, k_begin + c * GKC, smem + c * 2 * FCH);
      gemm32_stage<!TB>(D.B, (size_t)D.ldb, n0, k_begin + c * GKC, smem + (c * 2 + 1) * FCH);
    }
    for (int c = 0; c < nch; ++c) {
      if (c + 1 < nch)
        __builtin_amdgcn_s_waitcnt(wait_vm(4));  // chunk c+1's 4 loads per wave may fly on
      else
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
      gemm_barrier();
      const int cn = c + 2;
      gemm32_chunk<TA, TB>(acc, D, m0, n0, k_begin + cn * GKC, cn < nch, smem + (c % 3) * 2 * FCH,
                           smem + (cn % 3) * 2 * FCH);
    }
    gemm_barrier();
  }
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v) Cs[wn + bn * 16 + (lane & 15)][wm + bm * 16 + 4 * (lane >> 4) + v] = acc[bm][bn][v];
  __syncthreads();
  float* __restrict__ Cg = D.C;
  for (int idx = tid; idx < 4096; idx += 256) {
    const int r = idx & 63, c = idx >> 6;
    const size_t off = (size_t)(m0 + r) + (size_t)(n0 + c) * D.ldc;
    float v = D.alpha * Cs[c][r];
    if (D.beta != 0.f) v += D.beta * Cg[off];
    Cg[off] = v;
  }
}

static int gemm32_batched(DenseWork& w, const std::vector<Gemm32Desc>& in, bool ta, int flags, hipStream_t s) {
  std::vector<Gemm32Desc> descs;
  int total = 0;
  for (Gemm32Desc g : in) {
    if (g.mt <= 0 || g.nt <= 0) continue;
    if ((g.lda & 3) || (g.ldb & 3) || (reinterpret_cast<uintptr_t>(g.A) & 15) ||
        (reinterpret_cast<uintptr_t>(g.B) & 15)) {
      set_error("gemm32: operands must be 16-byte aligned with leading dimensions a multiple of 4");
      return RSD_ERR_ARG;
    }
    g.tile_base = total;
    total += (flags & kLowerC) ? g.mt * (g.mt + 1) / 2 : g.mt * g.nt;
    descs.push_back(g);
  }
  if (descs.empty()) return RSD_OK;
  Gemm32Args args{};
  args.nd = (int)descs.size();
  args.flags = flags;
  if (descs.size() <= (size_t)kInlineDesc) {
    for (size_t i = 0; i < descs.size(); ++i) args.d[i] = descs[i];
  } else {
    if (descs.size() > w.desc_cap) {
      set_error("gemm32: descriptor buffer too small");
      return RSD_ERR_CAPACITY;
    }
    static_assert(sizeof(Gemm32Desc) <= sizeof(GemmDesc), "FP32 descriptors fit the descriptor buffer");
    RSD_HIP(hipMemcpyAsync(w.d_desc, descs.data(), descs.size() * sizeof(Gemm32Desc), hipMemcpyHostToDevice, s));
    args.ext = reinterpret_cast<const Gemm32Desc*>(w.d_desc);
  }
  hipEvent_t pe = nullptr;
  if (prof_on()) {  // timed with the FP64 GEMMs' family; inner tiles as desc_ktiles counts them
    double kt = 0.0;
    for (const Gemm32Desc& g : descs) {
      const int ntile = (flags & kLowerC) ? g.mt : g.nt;
      for (int ti = 0; ti < g.mt; ++ti)
        for (int tj = 0; tj < ntile; ++tj) {
          if ((flags & kLowerC) && tj > ti) continue;
          int klo = 0, khi = g.kt;
          if (flags & kKloRow) klo = std::max(klo, ti);
          if (flags & kKloCol) klo = std::max(klo, tj);
          if (flags & kKhiRow) khi = std::min(khi, ti + 1);
          if (khi > klo) kt += khi - klo;
        }
    }
    pe = prof_begin(kProfGemm, kt * 2.0 * 64.0 * 64.0 * 64.0, 0.0, s);
  }
  if (ta)
    hipLaunchKernelGGL((gemm32_kernel<true, false>), dim3(total), dim3(256), 0, s, args);
  else
    hipLaunchKernelGGL((gemm32_kernel<false, false>), dim3(total), dim3(256), 0, s, args);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------------ 64x64 diagonal factor
// 1/sqrt(d) to full FP64 accuracy for the pivots, instead of an IEEE sqrt
// and a divide on the serial pivot chain: the hardware estimate y0
// (v_rsq_f64, relative error 5e-8 measured on gfx950) and one third-order
// correction y = y0 (1 + e/2 + 3e^2/8), e = 1 - d y0^2, whose error is
// O(e^3): within 2 ulp (max 2.2e-16 over 2^20 inputs spanning 2^-56..2^56).
__device__ __forceinline__ double rsqrt_nr(double d) {
  const double y0 = __builtin_amdgcn_rsq(d);
  const double e = fma(-d * y0, y0, 1.0);
  return fma(y0 * e, fma(0.375, e, 0.5), y0);
}

// Wave-uniform read of a double held by one lane (lane index known at
// compile time in the unrolled pivot loop).
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// D += P(16x16) * Q(16x16) on the matrix cores; P and Q point at the
// blocks' first elements in row-major LDS arrays with row strides PS, QS.
template <int PS, int QS>
__device__ __forceinline__ v4d mm16(const double* P, const double* Q, v4d acc, int lane) {
  const int r = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(P[r * PS + 4 * kb + kq], Q[(4 * kb + kq) * QS + r], acc, 0, 0, 0);
  return acc;
}

struct DiagSmem {
  double col[16][64];    // the owner wave's scaled columns, for its own deferred updates
  double P[2][64][18];   // finished 16-column panels, for the waves to their right
  double Ls[64][65];     // the factor, written panel by panel (upper triangle zero)
  double Xs[64][65];     // its inverse, lower 16x16 blocks
  double Tt[8][16][17];  // partial sums of the inverse's off-diagonal blocks
  double U[48][66];      // the previous panel's rank-16 term for columns 16..63, [column - 16][row]
  double ry[64];         // 1 / L_jj: the pivots' reciprocal square roots
  int bad;
};

// X_bb = inv(L_bb) for diagonal block b, by one wave: lane c < 16 solves
// column c by column-oriented forward substitution (16 steps on the chain,
// the row updates of a step independent of one another), with the pivots'
// reciprocals from the factorization instead of divides.  (Reading L_bb's
// columns ahead into registers is faster alone, but the walker is at its
// register limit here and then spills.)
__device__ __forceinline__ void diag_block_inverse(DiagSmem& sm, int b, int l) {
  if (l < 16) {
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = (i == l) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      x[k] *= sm.ry[16 * b + k];
#pragma unroll
      for (int i = k + 1; i < 16; ++i) x[i] -= sm.Ls[16 * b + i][16 * b + k] * x[k];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) sm.Xs[16 * b + i][16 * b + l] = x[i];
  }
}

// 16x16 blocks in MFMA layout (lane holds rows (lane>>4)+4v of column lane&15)
__device__ __forceinline__ v4d blk_load(const double (*T)[17], int lane) {
  v4d t;
#pragma unroll
  for (int v = 0; v < 4; ++v) t[v] = T[(lane >> 4) + 4 * v][lane & 15];
  return t;
}
template <int DS>
__device__ __forceinline__ void blk_store(double* dst, v4d t, double sign, int lane) {
#pragma unroll
  for (int v = 0; v < 4; ++v) dst[((lane >> 4) + 4 * v) * DS + (lane & 15)] = sign * t[v];
}
// a volatile int in LDS, addressed as such (a generic pointer would make the
// access a flat one, which also waits on the vector memory counter)
__device__ __forceinline__ volatile __attribute__((address_space(3))) int* lds_int(int* p) {
  return (volatile __attribute__((address_space(3))) int*)(p);
}
__device__ __forceinline__ volatile __attribute__((address_space(3))) double* lds_dbl(double* p) {
  return (volatile __attribute__((address_space(3))) double*)(p);
}
// a NaN with a payload no arithmetic produces: "not written yet"
constexpr double kNotYet = __builtin_bit_cast(double, 0x7ff4dead0000beefull);
__device__ __forceinline__ bool not_yet(double v) {
  return __double_as_longlong(v) == 0x7ff4dead0000beefll;
}
// 1/d to full FP64 accuracy: the hardware estimate and two Newton steps
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  return fma(fma(-d, r, 1.0), r, r);
}
// orders a wave's LDS writes before its later reads of other lanes' elements
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Cholesky of one 64x64 diagonal tile held in registers by 256 threads, and
// its triangular inverse: thread (w, l) has row l, columns 16w..16w+15 in
// a[].  The tile is done in four 16-column panels.  The wave that owns a
// panel factors it alone, as a software pipeline around the serial pivot
// chain (lane read of the pivot, rsqrt_nr, scale, lane read of the next
// column's entry): every other update of pivot q uses column entries read
// from LDS one iteration earlier.  The finished panel goes to LDS; after a
// barrier, its rank-16 term for all the columns to its right is formed by
// MFMA 16x16x16 products (the lower blocks only, spread over the waves) and
// handed over through LDS after a second barrier: each wave subtracts its
// own columns' term, and the owner's pivot loop carries no term of the
// previous panel.
// The inverse X = inv(L) is built from 16x16 blocks, X_bb = inv(L_bb) and
// X_bm = -X_bb (sum_{m'=m}^{b-1} L_bm' X_m'm), on the waves that the panels
// leave idle: the blocks of panels 0..2 while the later panels factor, and
// after the last panel two levels of MFMA 16x16x16 products over all four
// waves.
// On return (after a barrier) sm.Ls holds L (upper triangle zero), sm.Xs the
// lower blocks of X, and sm.bad the LAPACK-style INFO of the first
// non-positive pivot (k0 + j + 1) or 0x7fffffff.
// hook(b) runs while panel b+1 is factored (b = 0..2) on the waves that do
// not own it, and after the first level of the last products on every
// wave: points where a wave has slack.
template <class Hook>
__device__ __forceinline__ void diag_factor64(double (&a)[16], DiagSmem& sm, int k0, Hook&& hook) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  if (tid == 0) sm.bad = 0x7fffffff;
  const v4d z = {0.0, 0.0, 0.0, 0.0};
  auto Lb = [&](int i, int m) { return &sm.Ls[16 * i][16 * m]; };
  auto Xb = [&](int i, int m) { return &sm.Xs[16 * i][16 * m]; };
  auto Tb = [&](int s) { return &sm.Tt[s][0][0]; };
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    double(*P)[18] = sm.P[b & 1];
    double(*Pp)[18] = sm.P[(b + 1) & 1];  // the previous panel (b-1)
    if (b > 0) {
      // U = the previous panel's rank-16 term for the columns to its right:
      // block (rb, cb) = Pp[16rb.., :] Pp[16cb.., :]^T for rb >= cb >= b (the
      // blocks above the diagonal are not needed), by MFMA, spread over the
      // waves; every wave then subtracts its own columns' blocks
      auto u_block = [&](int rb, int cb) {
        const int r = l & 15, kq = l >> 4;
        v4d u = z;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
          u = __builtin_amdgcn_mfma_f64_16x16x4f64(Pp[16 * rb + r][4 * kb + kq], Pp[16 * cb + r][4 * kb + kq], u, 0,
                                                   0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) sm.U[16 * (cb - 1) + r][16 * rb + kq + 4 * v] = u[v];
      };
      if (b == 1) {
        if (w == 0) { u_block(1, 1); u_block(3, 3); }
        if (w == 1) { u_block(2, 1); u_block(2, 2); }
        if (w == 2) u_block(3, 1);
        if (w == 3) u_block(3, 2);
      }
      if (b == 2) {
        if (w == 0) u_block(2, 2);
        if (w == 1) u_block(3, 2);
        if (w == 2) u_block(3, 3);
      }
      if (b == 3 && w == 0) u_block(3, 3);
      if (b == 3 && w == 1) {
        // the last panel's columns, rows 48..63, marked as not yet written
        // (the wave that inverts L_33 behind the pivots watches them)
#pragma unroll
        for (int r = 0; r < 4; ++r) sm.col[4 * (l >> 4) + r][48 + (l & 15)] = kNotYet;
      }
      __syncthreads();
      if (w >= b) {
        // rows above the column block are above the diagonal: not needed
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] -= sm.U[16 * (w - 1) + q][l];
      }
      // the inverse's blocks that panels 0..b-1 allow, on the waves left of
      // this panel's owner (Tt slots: 0 L10 X00, 1 L21 X11, 2 L20 X00,
      // 3 L31 X11, 4 L30 X00)
      if (b == 1 && w == 0) diag_block_inverse(sm, 0, l);
      if (b == 2) {
        if (w == 1) diag_block_inverse(sm, 1, l);
        if (w == 0) blk_store<17>(Tb(0), mm16<65, 65>(Lb(1, 0), Xb(0, 0), z, l), 1.0, l);
      }
      if (b == 3) {
        if (w == 2) diag_block_inverse(sm, 2, l);
        if (w == 1) {
          blk_store<65>(Xb(1, 0), mm16<65, 17>(Xb(1, 1), Tb(0), z, l), -1.0, l);
          blk_store<17>(Tb(1), mm16<65, 65>(Lb(2, 1), Xb(1, 1), z, l), 1.0, l);
        }
        if (w == 0) {
          blk_store<17>(Tb(2), mm16<65, 65>(Lb(2, 0), Xb(0, 0), z, l), 1.0, l);
          blk_store<17>(Tb(3), mm16<65, 65>(Lb(3, 1), Xb(1, 1), z, l), 1.0, l);
          blk_store<17>(Tb(4), mm16<65, 65>(Lb(3, 0), Xb(0, 0), z, l), 1.0, l);
        }
      }
      if (w != b) hook(b - 1);
      if (b == 3 && w == 1) {
        // X_33 = inv(L_33) behind the last panel's pivots, by column-oriented
        // substitution as in diag_block_inverse, so that it is done when the
        // panel is.  Step k waits until column k's entries in rows 48..63 are
        // no longer marked (the owner writes a column with one LDS store),
        // and forms 1/L_kk itself.  Lane (g, c) holds rows 4g..4g+3 of column
        // c (few registers: the walker is near its limit here); step k's x_k
        // comes from lane (k/4, c) by a shuffle.
        const int c = l & 15, g = l >> 4;
        double x[4], d, e[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) x[r] = (4 * g + r == c) ? 1.0 : 0.0;
        // column k's diagonal entry and this lane's four entries, once written
        auto fetch = [&](int k) {
          bool pending;
          do {
            d = *lds_dbl(&sm.col[k][48 + k]);
            pending = not_yet(d);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              e[r] = *lds_dbl(&sm.col[k][48 + 4 * g + r]);
              pending |= not_yet(e[r]);
            }
            if (pending) __builtin_amdgcn_s_sleep(1);
          } while (pending);
        };
        fetch(0);
        double rk = recip(d);
        // step k; column k+1 is fetched and its reciprocal formed while step
        // k's shuffle is in flight
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const double xs = x[k & 3] * rk;
          if (g == (k >> 2)) x[k & 3] = xs;
          const double xk = __shfl(xs, 16 * (k >> 2) + c, 64);
          double ek[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) ek[r] = e[r];
          if (k + 1 < 16) {
            fetch(k + 1);
            rk = recip(d);
          }
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (4 * g + r > k) x[r] -= ek[r] * xk;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) sm.Xs[48 + 4 * g + r][48 + c] = x[r];
      }
    }
    if (w == b) {
      int first_bad = 64;
      double yq = 0.0, vprev = 0.0, rd[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) rd[q] = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int j = 16 * b + q;
        const double dj = readlane_d(a[q], j);
        first_bad = (dj > 0.0 || first_bad < 64) ? first_bad : j;
        const double y = rsqrt_nr(dj);
        yq = (l == q) ? y : yq;
        // lane j holds dj itself, so one product gives both L_jj and the column
        const double v = (l >= j) ? a[q] * y : 0.0;
        a[q] = v;
        sm.col[q][l] = v;
        if (q + 1 < 16) {
          // column q+1 before the next pivot: the previous pivot's term
          // (operands already in registers), then this pivot's term by a lane
          // read (the serial chain)
          if (q >= 1) a[q + 1] -= vprev * rd[q + 1];
          a[q + 1] -= v * readlane_d(v, j + 1);
        }
        // the previous pivot's terms for the columns right of q+1, from registers
        if (q >= 1) {
#pragma unroll
          for (int q2 = q + 2; q2 < 16; ++q2) a[q2] -= vprev * rd[q2];
        }
        // this pivot's column entries, read now and used one iteration later
#pragma unroll
        for (int q2 = q + 2; q2 < 16; ++q2) rd[q2] = sm.col[q][16 * b + q2];
        vprev = v;
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) P[l][q] = a[q];
#pragma unroll
      for (int q = 0; q < 16; ++q) sm.Ls[l][16 * b + q] = a[q];
      if (l < 16) sm.ry[16 * b + l] = yq;
      if (first_bad < 64 && l == 0) atomicMin(&sm.bad, k0 + first_bad + 1);
    }
    __syncthreads();
  }
  // after the last panel (X_33 is done by then): X_21, X_20, then X_32, X_31, X_30
  // (Tt slots: 5 L20 X00 + L21 X10, 6 L32 X22, 7 L30 X00 + L31 X10)
  if (w == 2) blk_store<65>(Xb(2, 1), mm16<65, 17>(Xb(2, 2), Tb(1), z, l), -1.0, l);
  if (w == 1) {
    blk_store<17>(Tb(5), mm16<65, 65>(Lb(2, 1), Xb(1, 0), blk_load(sm.Tt[2], l), l), 1.0, l);
    wave_lds_order();
    blk_store<65>(Xb(2, 0), mm16<65, 17>(Xb(2, 2), Tb(5), z, l), -1.0, l);
  }
  if (w == 0) {
    blk_store<17>(Tb(6), mm16<65, 65>(Lb(3, 2), Xb(2, 2), z, l), 1.0, l);
    blk_store<17>(Tb(7), mm16<65, 65>(Lb(3, 1), Xb(1, 0), blk_load(sm.Tt[4], l), l), 1.0, l);
  }
  hook(3);
  __syncthreads();
  if (w == 0) blk_store<65>(Xb(3, 2), mm16<65, 17>(Xb(3, 3), Tb(6), z, l), -1.0, l);
  if (w == 1) {
    blk_store<17>(Tb(3), mm16<65, 65>(Lb(3, 2), Xb(2, 1), blk_load(sm.Tt[3], l), l), 1.0, l);
    wave_lds_order();
    blk_store<65>(Xb(3, 1), mm16<65, 17>(Xb(3, 3), Tb(3), z, l), -1.0, l);
  }
  if (w == 2) {
    blk_store<17>(Tb(7), mm16<65, 65>(Lb(3, 2), Xb(2, 0), blk_load(sm.Tt[7], l), l), 1.0, l);
    wave_lds_order();
    blk_store<65>(Xb(3, 0), mm16<65, 17>(Xb(3, 3), Tb(7), z, l), -1.0, l);
  }
  __syncthreads();
}

// the inverse (column-major, zero upper blocks; written through) and the INFO word
__device__ __forceinline__ void diag_store_inverse(const DiagSmem& sm, double* __restrict__ dinv, int* __restrict__ info) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 16 * w + q;
    wt_store(dinv + (size_t)l + (size_t)c * 64, ((l >> 4) >= w) ? sm.Xs[l][c] : 0.0);
  }
  if (tid == 0 && sm.bad != 0x7fffffff) atomicMin(info, sm.bad);
}

// the factor into the matrix at (k0, k0)
__device__ __forceinline__ void diag_store_factor(const DiagSmem& sm, double* __restrict__ A, int ld, int k0) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = 16 * w + q;
    A[(size_t)(k0 + l) + (size_t)(k0 + c) * ld] = sm.Ls[l][c];
  }
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

int dense_info_reset(DenseWork& w, hipStream_t s) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, w.d_info, 0x7fffffff);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------- tile hand-off (shared by the
// dataflow Cholesky and the triangular solves)

// Spin (relaxed, agent scope) until *flag holds gen; a wait that runs out
// sets the error word and returns (the result is then garbage, never a hang).
// Once the error word is set every later wait returns at once, so a failed
// launch drains in one time-out rather than one per dependency.  The limit is
// wall-clock time, read from the GPU's constant 100 MHz counter: 20 s for any
// one wait, whatever the launch's size or what else shares the GPU.
constexpr unsigned long long kWaitTicks = 20ull * 100000000ull;

// rec (may be null): the first wait to time out records [1, who, what, the
// flag's value], for the host's error message.
__device__ __forceinline__ void wait_flag(const int* flag, int gen, int* err, int* rec = nullptr, int who = 0,
                                          int what = 0) {
  int it = 0;
  unsigned long long t0 = 0;
  while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen) {
    if ((++it & 1023) == 0) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      const unsigned long long now = __builtin_amdgcn_s_memrealtime();
      if (t0 == 0) {
        t0 = now;
      } else if (now - t0 > kWaitTicks) {
        if (rec && atomicCAS(&rec[0], 0, 1) == 0) {
          rec[1] = who;
          rec[2] = what;
          rec[3] = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        atomicOr(err, 1);
        break;
      }
    }
    __builtin_amdgcn_s_sleep(1);
  }
}


// --------------------------------------------- dataflow (tile) Cholesky
// Left-looking Cholesky of an n x n matrix (n = 64 nt) in one launch.
//
// Ticket 0 is the walker: one workgroup that runs the serial chain down the
// diagonal by itself.  For column j it takes the partial sum of tile (j,j)
// (all updates but the last), subtracts the last update L_{j,j-1} L_{j,j-1}^T
// from the tile it computed one column earlier and still holds in
// registers, factors the tile (diag_factor64, which also writes the 64x64
// inverse), publishes it, and then turns the partial sum of tile (j+1,j)
// into L_{j+1,j} with the inverse it has just built.  No hand-off between
// workgroups sits on the diagonal chain.
//
// Tickets 1.. are the lower tiles (i,j), one workgroup each: first the
// early tiles (c,c), (c+1,c), (c+2,c) of columns 0..kAhead-1, then for each
// column j the early tiles of column j + kAhead followed by column j's other
// tiles (rows j+3..).  The early tiles thus start kAhead columns before the
// walker needs them and take their terms as they are published, instead of
// running a long accumulation at the end: (c,c) and (c+1,c) are the walker's
// partial sums, and (c+2,c) feeds the last term of partial tile (c+2,c+1),
// which the walker needs one column later.  A task keeps its tile in MFMA
// accumulators and subtracts
// L_ik L_jk^T for every earlier column k as soon as tiles (i,k) and (j,k)
// are published.  Tiles (j,j) and (j+1,j) stop there and publish their
// partial sums for the walker (the diagonal one without the k = j-1 term);
// every other tile waits for the walker's diagonal inverse and solves
// L_ij = A_ij L_jj^-T as a product with it.
//
// Hand-off: the producer writes its tile through to the device's coherence
// point (wt_store), every wave waits for its stores, and after a barrier
// one thread stores the launch's generation into the tile's flag; the
// consumer spins on the flag and reads the tile with agent-scope loads (no
// cache-wide fence on either side).  The walker waits only for the two
// partial tiles of the column it is on.  Other tiles wait only for lower
// tickets or for the walker; early tiles wait for other tiles, some of
// higher tickets, but at any time only the early tiles within kAhead
// columns of the walker's (about 3 kAhead workgroups) can be waiting, so the
// other tiles always have workgroups to run on.  Every wait is bounded.
constexpr int kDagStage = 2 * 8192;  // two staged terms of two 64x64 tiles
constexpr int kAhead = 6;             // columns by which the early tiles start ahead
constexpr int kEarly = 3;             // early tiles per column: (c,c), (c+1,c), (c+2,c)
constexpr int kDiagDoubles = (sizeof(DiagSmem) + 15) / 16 * 2;  // the walker's share, 16-byte aligned
constexpr int kDagSmemDoubles = kDiagDoubles > kDagStage ? kDiagDoubles : kDagStage;
static_assert(kDagStage <= kDagSmemDoubles && 2 * 64 * 66 <= kDagSmemDoubles, "staging fits the shared buffer");

struct DagArgs {
  double* A;
  int ld, nt;
  double* dinv;  // nt inverted 64x64 diagonal factors, column-major
  double* part;  // 2 nt partial tiles, column-major 64x64: (j,j) at 2j, (j+1,j) at 2j+1;
                 // then (with X) the transposed diagonal blocks of X, ns tiles
  int* sync;     // [0] ticket, [1] error word, [2, 2+nt*nt) tile flags, then 2 nt partial
                 // flags, then (with X) ns*ns flags of X's tiles
  int gen;
  int* info;
  int* rec;      // the first timed-out wait: [1, ticket, flag index, flag value]
  // optional: X = inv(L_S) of the trailing block L_S (tiles s0..nt-1, ns = nt - s0
  // of them) into X (leading dimension ldx), formed in the same launch; s0 < 0: none
  int s0;
  double* X;
  int ldx;
};

// a 64x64 tile (column-major, leading dimension ld) <-> MFMA accumulator
// layout; Published: a tile written by another workgroup of the launch
// (agent-scope loads)
template <bool Published = false>
__device__ __forceinline__ void tile_load(v4d (&acc)[2][2], const double* src, size_t ld) {
  const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double* e = src + (size_t)(wm + 16 * bm + (lane >> 4) + 4 * v) + (size_t)(wn + 16 * bn + (lane & 15)) * ld;
        acc[bm][bn][v] = Published ? wt_load(e) : *e;
      }
}

// a published 64x64 tile (column-major, leading dimension 64) as the A
// operands of a 64-deep MFMA product: a0[st] = P(wm + m, 4 st + kq),
// a1[st] = P(wm + 16 + m, 4 st + kq), m = lane & 15, kq = lane >> 4
__device__ __forceinline__ void tile_load_a(double (&a0)[16], double (&a1)[16], const double* src) {
  const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, m = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    a0[st] = wt_load(src + (wm + m) + (4 * st + kq) * 64);
    a1[st] = wt_load(src + (wm + 16 + m) + (4 * st + kq) * 64);
  }
}

__device__ __forceinline__ void tile_store_direct(const v4d (&acc)[2][2], double* dst, size_t ld) {
  const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        wt_store(dst + (size_t)(wm + 16 * bm + (lane >> 4) + 4 * v) + (size_t)(wn + 16 * bn + (lane & 15)) * ld,
                 acc[bm][bn][v]);
}

// the transpose of a tile held in accumulator layout, written through
__device__ __forceinline__ void tile_store_direct_t(const v4d (&acc)[2][2], double* dst, size_t ld) {
  const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        wt_store(dst + (size_t)(wn + 16 * bn + (lane & 15)) + (size_t)(wm + 16 * bm + (lane >> 4) + 4 * v) * ld,
                 acc[bm][bn][v]);
}

// One term of a tile's accumulation: tiles (i,k) and (j,k) (one tile when
// i == j) staged in LDS, (i,k) first.
__device__ __forceinline__ void stage_term(const double* __restrict__ A, size_t ld, int i, int j, int k,
                                           double* __restrict__ dst) {
  stage_tile(A + 64 * i + (size_t)(64 * k) * ld, ld, dst);
  if (i != j) stage_tile(A + 64 * j + (size_t)(64 * k) * ld, ld, dst + 4096);
}

// Two 64x64 tiles from anywhere (the first at dst, the second at dst + 4096).
__device__ __forceinline__ void stage_pair(const double* p1, size_t ld1, const double* p2, size_t ld2,
                                           double* __restrict__ dst) {
  stage_tile(p1, ld1, dst);
  stage_tile(p2, ld2, dst + 4096);
}

// nacc += L_ik L_jk^T (nacc: the tile negated) from a staged term (cur), the
// next term's loads (into nxt, when pre) in flight meanwhile.  All 16 steps'
// operands are read into registers first, so that the LDS reads run ahead of
// the MFMAs that use them instead of each step waiting for its own reads,
// and so that they precede the direct-to-LDS loads of the next term: LDS
// reads issued after those loads would wait for them to land.
__device__ __forceinline__ void term_product(v4d (&nacc)[2][2], const double* __restrict__ A, size_t ld, int i, int j,
                                             int knext, bool pre, const double* __restrict__ cur,
                                             double* __restrict__ nxt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const double* cb = cur + (i == j ? 0 : 4096);
  const int m = lane & 15, kq = lane >> 4;
  double a0[16], a1[16], b0[16], b1[16];
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int kr = 4 * st + kq;
    a0[st] = cur[staged(wm + m, kr)];
    a1[st] = cur[staged(wm + 16 + m, kr)];
    b0[st] = cb[staged(wn + m, kr)];
    b1[st] = cb[staged(wn + 16 + m, kr)];
  }
  if (pre) stage_term(A, ld, i, j, knext, nxt);
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    nacc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b0[st], nacc[0][0], 0, 0, 0);
    nacc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b1[st], nacc[0][1], 0, 0, 0);
    nacc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b0[st], nacc[1][0], 0, 0, 0);
    nacc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b1[st], nacc[1][1], 0, 0, 0);
  }
}

__device__ __forceinline__ void tile_negate(v4d (&acc)[2][2]) {
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = -acc[bm][bn];
}

// acc -= sum_{k < kend} L_ik L_jk^T, each term once tiles (i,k), (j,k) are
// published.  Terms go through two LDS buffers of 2 x 64 x 64 doubles: the
// loads of term k+1 overlap the products of term k whenever term k+1 is
// already published when term k starts.
__device__ __forceinline__ void tile_accumulate(v4d (&acc)[2][2], const double* __restrict__ A, int ld, int nt, int i,
                                                int j, int kend, const int* flags, int gen, int* err, double* smem,
                                                int* rec, int who) {
  const int tid = threadIdx.x;
  __shared__ int kready_s;
  int kready = 0;
  // wait for term k, then take every later term that is already published
  // too: one barrier per batch, not per term
  auto ready = [&](int k) {
    if (tid == 0) {
      wait_flag(&flags[i + nt * k], gen, err, rec, who, i + nt * k);
      if (i != j) wait_flag(&flags[j + nt * k], gen, err, rec, who, j + nt * k);
      int kk = k + 1;
      while (kk < kend && __hip_atomic_load(&flags[i + nt * kk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen &&
             __hip_atomic_load(&flags[j + nt * kk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
        ++kk;
      kready_s = kk;
    }
    __syncthreads();
    kready = kready_s;
    wt_seen();
  };
  if (kend <= 0) return;
  tile_negate(acc);  // the terms add to the negated tile (exact: negation commutes with rounding)
  ready(0);
  stage_term(A, (size_t)ld, i, j, 0, smem);
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
  for (int k = 0; k < kend; ++k) {
    double* cur = smem + (k & 1) * 8192;
    double* nxt = smem + ((k + 1) & 1) * 8192;
    const bool more = k + 1 < kend, pre = more && k + 1 < kready;
    term_product(acc, A, (size_t)ld, i, j, k + 1, pre, cur, nxt);
    if (more && !pre) {
      ready(k + 1);
      stage_term(A, (size_t)ld, i, j, k + 1, nxt);
    }
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
  }
  tile_negate(acc);
}

// nacc += T1 T2^T for a staged pair (T1 at cur, T2 at cur + 4096); stage()
// issues the next pair's loads into nxt (when pre), after this pair's
// operands are read (as in term_product).
template <class Stage>
__device__ __forceinline__ void pair_product(v4d (&nacc)[2][2], bool pre, const double* __restrict__ cur,
                                             Stage&& stage) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const double* cb = cur + 4096;
  const int m = lane & 15, kq = lane >> 4;
  double a0[16], a1[16], b0[16], b1[16];
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int kr = 4 * st + kq;
    a0[st] = cur[staged(wm + m, kr)];
    a1[st] = cur[staged(wm + 16 + m, kr)];
    b0[st] = cb[staged(wn + m, kr)];
    b1[st] = cb[staged(wn + 16 + m, kr)];
  }
  if (pre) stage();
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    nacc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b0[st], nacc[0][0], 0, 0, 0);
    nacc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b1[st], nacc[0][1], 0, 0, 0);
    nacc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b0[st], nacc[1][0], 0, 0, 0);
    nacc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b1[st], nacc[1][1], 0, 0, 0);
  }
}

// acc -= sum_{kbeg <= k < kend} T1_k T2_k^T, with src(k, p1, ld1, p2, ld2, f1,
// f2) naming term k's two tiles and their flags (indices into flags); each
// term once both flags are published (as tile_accumulate).
template <class Src>
__device__ __forceinline__ void pair_accumulate(v4d (&acc)[2][2], int kbeg, int kend, Src&& src, const int* flags,
                                                int gen, int* err, double* smem, int* rec, int who) {
  const int tid = threadIdx.x;
  __shared__ int kready_x;
  int kready = kbeg;
  auto published = [&](int k) {
    const double *p1, *p2;
    size_t l1, l2;
    int f1, f2;
    src(k, p1, l1, p2, l2, f1, f2);
    return __hip_atomic_load(&flags[f1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen &&
           __hip_atomic_load(&flags[f2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen;
  };
  auto ready = [&](int k) {
    if (tid == 0) {
      const double *p1, *p2;
      size_t l1, l2;
      int f1, f2;
      src(k, p1, l1, p2, l2, f1, f2);
      wait_flag(&flags[f1], gen, err, rec, who, f1);
      wait_flag(&flags[f2], gen, err, rec, who, f2);
      int kk = k + 1;
      while (kk < kend && published(kk)) ++kk;
      kready_x = kk;
    }
    __syncthreads();
    kready = kready_x;
    wt_seen();
  };
  auto stage = [&](int k, double* dst) {
    const double *p1, *p2;
    size_t l1, l2;
    int f1, f2;
    src(k, p1, l1, p2, l2, f1, f2);
    stage_pair(p1, l1, p2, l2, dst);
  };
  if (kend <= kbeg) return;
  tile_negate(acc);
  ready(kbeg);
  stage(kbeg, smem);
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
  for (int k = kbeg; k < kend; ++k) {
    double* cur = smem + ((k - kbeg) & 1) * 8192;
    double* nxt = smem + ((k - kbeg + 1) & 1) * 8192;
    const bool more = k + 1 < kend, pre = more && k + 1 < kready;
    pair_product(acc, pre, cur, [&] { stage(k + 1, nxt); });
    if (more && !pre) {
      ready(k + 1);
      stage(k + 1, nxt);
    }
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
  }
  tile_negate(acc);
}

// out = acc Dinv^T, Dinv a column-major 64x64 tile in global memory
__device__ __forceinline__ void tile_trsm(const v4d (&acc)[2][2], const double* __restrict__ Dg, double* smem,
                                          v4d (&out)[2][2]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  // Cs[kk][m] = acc(m, kk); Dinv staged, so that element (n, kk) = Dinv^T(kk, n)
  double(*Cs)[66] = reinterpret_cast<double(*)[66]>(smem);
  double* Dj = smem + 64 * 66;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v) Cs[wn + 16 * bn + (lane & 15)][wm + 16 * bm + (lane >> 4) + 4 * v] = acc[bm][bn][v];
  stage_tile(Dg, 64, Dj);
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) out[bm][bn] = (v4d){0.0, 0.0, 0.0, 0.0};
  double a0[16], a1[16], b0[16], b1[16];  // every step's operands read ahead of the MFMAs
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int kr = 4 * st + (lane >> 4);
    a0[st] = Cs[kr][wm + (lane & 15)];
    a1[st] = Cs[kr][wm + 16 + (lane & 15)];
    b0[st] = Dj[staged(wn + (lane & 15), kr)];
    b1[st] = Dj[staged(wn + 16 + (lane & 15), kr)];
  }
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    out[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b0[st], out[0][0], 0, 0, 0);
    out[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b1[st], out[0][1], 0, 0, 0);
    out[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b0[st], out[1][0], 0, 0, 0);
    out[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b1[st], out[1][1], 0, 0, 0);
  }
  __syncthreads();
}

// tile store through LDS, so each column is written contiguously; written
// through, every wave's stores complete at the closing barrier
__device__ __forceinline__ void tile_store(const v4d (&out)[2][2], double* dst, int ld, double* smem) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  double(*Cs)[66] = reinterpret_cast<double(*)[66]>(smem);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v) Cs[wn + 16 * bn + (lane & 15)][wm + 16 * bm + (lane >> 4) + 4 * v] = out[bm][bn][v];
  __syncthreads();
  for (int idx = tid; idx < 4096; idx += 256) {
    const int r = idx & 63, c = idx >> 6;
    wt_store(dst + (size_t)r + (size_t)c * ld, Cs[c][r]);
  }
  wt_stores_done();
  __syncthreads();
}

// dst = X (a 64x64 tile, leading dimension ld) and acc -= X X^T, for X held
// in accumulator layout (column blocks cb0, cb1 of this wave's rows).  X goes to LDS transposed, T[kk][m] = X(m, kk), so
// that each of its columns is contiguous there: the stores (written through)
// leave from LDS as whole columns, and the products read the same copy.
// Every wave's stores are complete at the closing barrier.
__device__ __forceinline__ void tile_store_syrk(v4d (&acc)[2][2], const v4d (&X)[2][2], int cb0, int cb1,
                                                double* __restrict__ dst, size_t ld, double* smem) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  double(*T)[66] = reinterpret_cast<double(*)[66]>(smem);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        T[16 * (bn ? cb1 : cb0) + (lane & 15)][wm + 16 * bm + (lane >> 4) + 4 * v] = X[bm][bn][v];
  __syncthreads();
  // 64 columns of 32 row pairs: 32 consecutive threads store one column
#pragma unroll
  for (int p = tid; p < 2048; p += 256) {
    const int c = p >> 5, r = 2 * (p & 31);
    const double2 x = *reinterpret_cast<const double2*>(&T[c][r]);
    wt_store(dst + (size_t)r + (size_t)c * ld, x.x);
    wt_store(dst + (size_t)r + 1 + (size_t)c * ld, x.y);
  }
  double a0[16], a1[16], b0[16], b1[16];  // every step's operands read ahead of the MFMAs
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int kr = 4 * st + (lane >> 4);
    a0[st] = T[kr][wm + (lane & 15)];
    a1[st] = T[kr][wm + 16 + (lane & 15)];
    b0[st] = T[kr][wn + (lane & 15)];
    b1[st] = T[kr][wn + 16 + (lane & 15)];
  }
  tile_negate(acc);  // adds the products to the negated tile (exact)
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b0[st], acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b1[st], acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b0[st], acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b1[st], acc[1][1], 0, 0, 0);
  }
  tile_negate(acc);
  wt_stores_done();
  __syncthreads();
}

// The walker's L_{j+1,j} = P X^T for output column blocks C0 and C1 of this
// wave's 32 rows (A operands a0, a1 of P in registers, X = inv(L_jj) in
// sm.Xs): column block c takes k-steps 0..4c+3 only, X^T being zero below
// its block diagonal (and zero inside the diagonal blocks below the diagonal).
template <int C0, int C1>
__device__ __forceinline__ void walker_lt_product(v4d (&Lt)[2][2], const double (&a0)[16], const double (&a1)[16],
                                                  const DiagSmem& sm, int lane) {
  constexpr int S0 = 4 * (C0 + 1), S1 = 4 * (C1 + 1);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) Lt[bm][bn] = (v4d){0.0, 0.0, 0.0, 0.0};
  double b0[S0], b1[S1];  // every step's operands read ahead of the MFMAs
  const int n0c = 16 * C0 + (lane & 15), n1c = 16 * C1 + (lane & 15);
#pragma unroll
  for (int st = 0; st < S1; ++st) {
    const int kr = 4 * st + (lane >> 4);
    if (st < S0) b0[st < S0 ? st : 0] = sm.Xs[n0c][kr];
    b1[st] = sm.Xs[n1c][kr];
  }
#pragma unroll
  for (int st = 0; st < S1; ++st) {
    if (st < S0) {
      Lt[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b0[st < S0 ? st : 0], Lt[0][0], 0, 0, 0);
      Lt[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b0[st < S0 ? st : 0], Lt[1][0], 0, 0, 0);
    }
    Lt[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[st], b1[st], Lt[0][1], 0, 0, 0);
    Lt[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[st], b1[st], Lt[1][1], 0, 0, 0);
  }
}

// One task of the dataflow Cholesky: ticket 0 the walker, the others a tile
// (see above).  Returns when the task is done.
__device__ __forceinline__ void dag_task(const DagArgs& g, int t, double* __restrict__ smem) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int nt = g.nt, ld = g.ld;
  double* __restrict__ A = g.A;
  int* err = g.sync + 1;
  int* flags = g.sync + 2;
  int* pflags = g.sync + 2 + nt * nt;
  int* rec = g.rec;  // the first timed-out wait
  v4d acc[2][2];

  if (t == 0) {  // the walker
    // acc: partial sum of the current diagonal tile, already including the
    // last update (the previous column's L_{j,j-1} L_{j,j-1}^T)
    DiagSmem& sm = *reinterpret_cast<DiagSmem*>(smem);
    if (tid == 0) wait_flag(&pflags[0], g.gen, err, rec, 0, nt * nt);
    __syncthreads();
    wt_seen();
    tile_load<true>(acc, g.part, 64);
    for (int j = 0; j < nt; ++j) {
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            sm.Ls[wm + 16 * bm + (lane >> 4) + 4 * v][wn + 16 * bn + (lane & 15)] = acc[bm][bn][v];
      __syncthreads();
      double a[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) a[q] = sm.Ls[lane][16 * wave + q];
      __syncthreads();
      // column j+1's partial tiles, (j+1,j) into sa0/sa1 (the A operands of
      // L_{j+1,j}'s product) and (j+1,j+1) into acc: each wave issues its
      // share of the loads at the first point where both are published (the
      // column's start, or a point of slack in the factorization), so that
      // they land during it; a wave that has not issued them by the end of
      // the factorization waits for the tiles there.  The loads go to
      // registers: loads into LDS would hold up every barrier until they land.
      double sa0[16], sa1[16];
      const double* sub_src = g.part + (size_t)(2 * j + 1) * 4096;
      const double* acc_src = g.part + (size_t)(2 * (j + 1)) * 4096;
      bool issued = j + 1 >= nt;
      auto issue = [&](bool wait) {
        if (issued) return;
        if (wait) {
          if (lane == 0) {
            wait_flag(&pflags[2 * j + 1], g.gen, err, rec, 0, nt * nt + 2 * j + 1);
            wait_flag(&pflags[2 * (j + 1)], g.gen, err, rec, 0, nt * nt + 2 * (j + 1));
          }
        } else {
          const int ready =
              __hip_atomic_load(&pflags[2 * j + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g.gen &&
              __hip_atomic_load(&pflags[2 * (j + 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g.gen;
          if (!__builtin_amdgcn_readfirstlane(ready)) return;
        }
        wt_seen();
        tile_load_a(sa0, sa1, sub_src);
        tile_load<true>(acc, acc_src, 64);
        issued = true;
      };
      issue(false);
      // the previous column's L_{j,j-1}: published by the last wave, which
      // waits for the first panel's pivots anyway (its stores are behind the
      // barriers above)
      if (j > 0 && tid == 192) wt_publish(&flags[j + nt * (j - 1)], g.gen);
      diag_factor64(a, sm, 64 * j, [&](int) { issue(false); });
      diag_store_inverse(sm, g.dinv + (size_t)j * 4096, g.info);
      wt_stores_done();
      __syncthreads();
      if (tid == 0) wt_publish(&flags[j + nt * j], g.gen);
      diag_store_factor(sm, A, ld, 64 * j);  // read by nobody in this launch
      if (j + 1 == nt) break;
      issue(true);
      __builtin_amdgcn_s_waitcnt(kWaitVm0);
      __syncthreads();
      // L_{j+1,j} = P_{j+1,j} X^T with X = inv(L_jj) still in sm.Xs (lower
      // blocks).  X^T is block upper triangular, so output column block c
      // takes only the k-steps of blocks 0..c; waves 0,1 form column blocks
      // 0 and 3, waves 2,3 blocks 1 and 2: 40 MFMAs each instead of 64
      v4d Lt[2][2];
      if (wave >> 1)
        walker_lt_product<1, 2>(Lt, sa0, sa1, sm, lane);
      else
        walker_lt_product<0, 3>(Lt, sa0, sa1, sm, lane);
      __syncthreads();
      // L_{j+1,j} out (written through), and the last update of the next
      // diagonal tile, the stores' completion waited for at the end
      const int cb0 = (wave >> 1) ? 1 : 0, cb1 = (wave >> 1) ? 2 : 3;
      tile_store_syrk(acc, Lt, cb0, cb1, A + 64 * (j + 1) + (size_t)(64 * j) * ld, ld, smem);
      // L_{j+1,j} is published at the start of the next column by a wave
      // that is idle there
    }
    return;
  }

  // tickets past the factorization's: X = inv(L_S) (see potrf_lower_x)
  if (t > nt * (nt + 1) / 2) {
    const int s0 = g.s0, ns = nt - s0, r = t - nt * (nt + 1) / 2 - 1;
    double* xdT = g.part + (size_t)2 * nt * 4096;   // X_jj^T = Dinv^T of the diagonal blocks
    const int xf = nt * nt + 2 * nt;                 // X's flags, after the partial flags
    if (r < ns) {
      // diagonal block j: X_jj = Dinv_J into X (read by the LL^T product after
      // the launch) and its transpose (the first term's operand of block
      // column j)
      const int jj = r, J = s0 + jj;
      if (tid == 0) wait_flag(&flags[J + nt * J], g.gen, err, rec, t, J + nt * J);
      __syncthreads();
      wt_seen();
      tile_load<true>(acc, g.dinv + (size_t)J * 4096, 64);
      tile_store_direct(acc, g.X + 64 * jj + (size_t)(64 * jj) * g.ldx, g.ldx);
      tile_store_direct_t(acc, xdT + (size_t)jj * 4096, 64);
      wt_stores_done();
      __syncthreads();
      if (tid == 0) wt_publish(&flags[xf + jj + ns * jj], g.gen);
      return;
    }
    // tile (xi, xj), xi > xj, in row order: X_ij = -Dinv_I sum_{k=j}^{i-1} L_Ik X_kj
    // (I = s0 + i).  The sum is accumulated transposed, acc = -sum X_kj^T L_Ik^T,
    // from X_kj^T (kept in X's upper block (j,k); X_jj^T in xdT), so that the
    // solve with Dinv_I is tile_trsm's product: X_ij^T = acc Dinv_I^T.
    int q = r - ns, xi = 1;
    while (q >= xi) {
      q -= xi;
      ++xi;
    }
    const int xj = q, I = s0 + xi;
    double* X = g.X;
    const size_t ldx = (size_t)g.ldx;
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn) acc[bm][bn] = (v4d){0.0, 0.0, 0.0, 0.0};
    pair_accumulate(
        acc, xj, xi,
        [&](int k, const double*& p1, size_t& l1, const double*& p2, size_t& l2, int& f1, int& f2) {
          if (k == xj) {
            p1 = xdT + (size_t)xj * 4096;
            l1 = 64;
          } else {
            p1 = X + 64 * xj + (size_t)(64 * k) * ldx;
            l1 = ldx;
          }
          f1 = xf + k + ns * xj;
          p2 = A + 64 * I + (size_t)(64 * (s0 + k)) * ld;
          l2 = (size_t)ld;
          f2 = I + nt * (s0 + k);
        },
        flags, g.gen, err, smem, rec, t);
    if (tid == 0) wait_flag(&flags[I + nt * I], g.gen, err, rec, t, I + nt * I);
    __syncthreads();
    wt_seen();
    v4d out[2][2];
    tile_trsm(acc, g.dinv + (size_t)I * 4096, smem, out);
    // X_ij^T into the upper block (j,i) for the later tiles of block column j;
    // X_ij into the lower block (i,j)
    tile_store_direct(out, X + 64 * xj + (size_t)(64 * xi) * ldx, ldx);
    tile_store_direct_t(out, X + 64 * xi + (size_t)(64 * xj) * ldx, ldx);
    wt_stores_done();
    __syncthreads();
    if (tid == 0) wt_publish(&flags[xf + xi + ns * xj], g.gen);
    return;
  }

  // the tile of ticket t (see above)
  int i, j;
  {
    auto esize = [nt](int c) { return c < nt ? min(kEarly, nt - c) : 0; };  // early tiles of column c
    int r = t - 1, c0 = 0;
    for (; c0 < kAhead && c0 < nt; ++c0) {
      if (r < esize(c0)) break;
      r -= esize(c0);
    }
    if (c0 < kAhead && c0 < nt) {
      j = c0;
      i = c0 + r;
    } else {
      for (j = 0;; ++j) {
        const int e = esize(j + kAhead), o = nt - j - kEarly > 0 ? nt - j - kEarly : 0;
        if (r < e) {
          i = j + kAhead + r;
          j = j + kAhead;
          break;
        }
        r -= e;
        if (r < o) {
          i = j + kEarly + r;
          break;
        }
        r -= o;
      }
    }
  }
  tile_load(acc, A + 64 * i + (size_t)(64 * j) * ld, ld);
  if (i == j || i == j + 1) {
    // partial sums for the walker: (j,j) without its last term, (j+1,j) complete
    tile_accumulate(acc, A, ld, nt, i, j, i == j ? max(j - 1, 0) : j, flags, g.gen, err, smem, rec, t);
    tile_store_direct(acc, g.part + (size_t)(2 * j + (i - j)) * 4096, 64);
    wt_stores_done();
    __syncthreads();
    if (tid == 0) wt_publish(&pflags[2 * j + (i - j)], g.gen);
    return;
  }
  tile_accumulate(acc, A, ld, nt, i, j, j, flags, g.gen, err, smem, rec, t);
  if (tid == 0) wait_flag(&flags[j + nt * j], g.gen, err, rec, t, j + nt * j);
  __syncthreads();
  wt_seen();
  v4d out[2][2];
  tile_trsm(acc, g.dinv + (size_t)j * 4096, smem, out);
  tile_store(out, A + 64 * i + (size_t)(64 * j) * ld, ld, smem);
  if (tid == 0) wt_publish(&flags[i + nt * j], g.gen);
}

// The launch is persistent: as many workgroups as the GPU holds at once
// (one per CU: the tile staging fills the LDS), each taking tickets in turn
// until they run out.  A task waits only for tasks of lower tickets, for the
// walker, or (the early tiles, at most kEarly kAhead of them) for tiles of
// higher tickets; since every ticket taken is held by a running workgroup,
// and a workgroup takes the next ticket as soon as it is free, the tasks
// waited for always run.  (With one workgroup per task, the hardware assigns
// the workgroups to the XCDs in advance: when the only XCD with workgroups
// left to start is full of waiting ones, the lowest untaken ticket never
// starts.)
__global__ __launch_bounds__(256) void potrf_dag_kernel(DagArgs g) {
  __shared__ __attribute__((aligned(16))) double smem[kDagSmemDoubles];
  __shared__ int t_s;
  const int ns = g.s0 >= 0 ? g.nt - g.s0 : 0;
  const int total = g.nt * (g.nt + 1) / 2 + ns * (ns + 1) / 2;  // the last ticket
  for (;;) {
    if (threadIdx.x == 0) {
      const int t = atomicAdd(&g.sync[0], 1);
      t_s = t;
      // each workgroup ends with one take past the last ticket: the last of
      // those clears the counter for the next launch
      if (t == total + (int)gridDim.x)
        __hip_atomic_store(&g.sync[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int t = t_s;
    __syncthreads();
    if (t > total) break;
    dag_task(g, t, smem);
    __syncthreads();
  }
}

// one persistent workgroup per CU
static int dag_workgroups() {
  static int n = 0;
  if (n == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      n = cus;
    else
      n = 256;
  }
  return n;
}

__global__ void dag_wait_reset_kernel(int* err, int* rec) {
  if (threadIdx.x == 0) *err = 0;
  if (threadIdx.x < 4) rec[threadIdx.x] = 0;
}

static int dag_reserve(DenseWork& w, int nt, hipStream_t s) {
  // The ticket and the error word have fixed places; the flags hold the
  // generation of the launch that set them, so launches of another order
  // (whose flags sit elsewhere) need no clearing in between.  Room for X's
  // flags and transposed diagonal blocks (up to nt of each side) is always
  // kept.
  const size_t need = 2 + (size_t)nt * nt + 2 * (size_t)nt + (size_t)nt * nt;
  if (need > w.dag_cap) {
    if (w.d_dag) RSD_HIP(hipFree(w.d_dag));
    RSD_HIP(hipMalloc(&w.d_dag, need * sizeof(int)));
    RSD_HIP(hipMemsetAsync(w.d_dag, 0, need * sizeof(int), s));
    if (!w.d_rec) {
      RSD_HIP(hipMalloc(&w.d_rec, 4 * sizeof(int)));
      RSD_HIP(hipMemsetAsync(w.d_rec, 0, 4 * sizeof(int), s));
    }
    w.dag_cap = need;
    w.dag_gen = 0;
  }
  const size_t need_part = 3 * (size_t)nt * 4096;
  if (need_part > w.part_cap) {
    if (w.d_part) RSD_HIP(hipFree(w.d_part));
    RSD_HIP(hipMalloc(&w.d_part, need_part * sizeof(double)));
    w.part_cap = need_part;
  }
  return RSD_OK;
}

int potrf_lower(DenseWork& w, double* A, int n, int ld, hipStream_t s) {
  return potrf_lower_x(w, A, n, ld, -1, nullptr, 0, s);
}

int potrf_lower_x(DenseWork& w, double* A, int n, int ld, int s0, double* X, int ldx, hipStream_t s) {
  const int nt = n / 64;
  if (nt < 1) return RSD_OK;
  if ((ld & 1) || (reinterpret_cast<uintptr_t>(A) & 15)) {
    // the tile loads move 16-byte pairs of rows
    set_error("potrf_lower: the matrix must be 16-byte aligned with an even leading dimension");
    return RSD_ERR_ARG;
  }
  if (X && (s0 < 0 || s0 >= nt || (ldx & 1) || ldx < 64 * (nt - s0) || (reinterpret_cast<uintptr_t>(X) & 15))) {
    set_error("potrf_lower_x: X needs 0 <= s0 < n/64, 16-byte alignment and an even ldx >= n - 64 s0");
    return RSD_ERR_ARG;
  }
  if (!X) s0 = -1;
  const int rc = dag_reserve(w, nt, s);
  if (rc) return rc;
  if (++w.dag_gen == 0x7fffffff) w.dag_gen = 1;
  DagArgs g{A, ld, nt, w.dinv, w.d_part, w.d_dag, w.dag_gen, w.d_info, w.d_rec, s0, X, ldx};
  const int ns = s0 >= 0 ? nt - s0 : 0;
  // algorithmic flops: n^3/3 for the factor, plus m^3/3 for the triangular inverse of its trailing
  // m = 64 ns rows when X is asked for
  const double nd = (double)n, md = 64.0 * ns;
  hipEvent_t pe = prof_begin(kProfPotrfDiag, (nd * nd * nd + md * md * md) / 3.0, 0.0, s);
  hipLaunchKernelGGL(potrf_dag_kernel, dim3(std::min(1 + nt * (nt + 1) / 2 + ns * (ns + 1) / 2, dag_workgroups())),
                     dim3(256), 0, s, g);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// ------------------------------------------- triangular solves (potrs)
// (L L^T) X = B for up to two right-hand sides, one persistent launch per
// direction.  Workgroup t takes block row k by an atomic ticket (forward:
// k = t, backward: k = nb-1-t), accumulates L_kj y_j for the blocks j it
// depends on as soon as each is published, applies the 64x64 diagonal
// inverse and publishes its own block: the 64 values written through, then
// (once every wave's stores are done) a store of the solve's generation into
// the block's flag; readers take published blocks with agent-scope loads.
// A workgroup only ever waits for blocks of lower tickets, which running
// workgroups hold, so the
// solve finishes whatever the residency; every wait is bounded and a wait
// that runs out sets the error word instead of hanging.
// Forward: y_k = Dinv_k (b_k - sum_{j<k} L_kj y_j).  Lane = row of block k,
// wave wv takes columns 16wv..16wv+15 of every L_kj; each wave waits for
// block j on its own (no workgroup barrier in the loop) after issuing the
// loads of L_kj, so the loads overlap the wait.
template <int NR>
__global__ __launch_bounds__(256) void potrs_fwd_kernel(const double* __restrict__ L, int ld,
                                                        const double* __restrict__ dinv, double* X, int ldx, int nb,
                                                        int* sync, int gen) {
  __shared__ int k_s;
  __shared__ double yv[NR][64];
  __shared__ double red[NR][4][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int* err = sync + 2;
  int* flags = sync + 3;
  if (tid == 0) {
    const int t = atomicAdd(&sync[0], 1);
    k_s = t;
    // every workgroup has its ticket: the last one clears the counter for the next launch
    if (t == nb - 1) __hip_atomic_store(&sync[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int k = k_s;
  const size_t row = (size_t)64 * k + lane;
  // Dinv_k and b_k depend on no other workgroup (only this one writes block
  // row k of X): load them now, off the chain of waits below
  const double* D = dinv + (size_t)k * 4096;
  double dv[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) dv[c] = D[lane + 64 * (16 * wv + c)];
  const double bk = wv < NR ? X[(size_t)wv * ldx + row] : 0.0;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  for (int j = 0; j < k; ++j) {
    const double* Lb = L + row + (size_t)(64 * j + 16 * wv) * ld;
    double lv[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) lv[c] = Lb[(size_t)c * ld];
    if (lane == 0) wait_flag(&flags[j], gen, err);
    __builtin_amdgcn_wave_barrier();
    wt_seen();
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double* yj = X + (size_t)r * ldx + (size_t)64 * j + 16 * wv;
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[r] += lv[c] * wt_load(yj + c);
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) red[r][wv][lane] = acc[r];
  __syncthreads();
  // t = b_k - sum; y_k = Dinv_k t
  if (wv < NR) yv[wv][lane] = bk - ((red[wv][0][lane] + red[wv][1][lane]) + (red[wv][2][lane] + red[wv][3][lane]));
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 16; ++c) t += dv[c] * yv[r][16 * wv + c];
    acc[r] = t;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; ++r) red[r][wv][lane] = acc[r];
  __syncthreads();
  if (wv < NR)
    wt_store(X + (size_t)wv * ldx + row, (red[wv][0][lane] + red[wv][1][lane]) + (red[wv][2][lane] + red[wv][3][lane]));
  wt_stores_done();
  __syncthreads();
  if (tid == 0) wt_publish(&flags[k], gen);
}

// Sum over the 64 lanes of a wave of p[c], c < 16 (this lane's shares of
// 16 column sums): the set is halved across lane bits 32, 16, 8, 4 (each lane
// keeps half and adds its partner's half), then added across bits 2 and 1.
// Lanes with (lane & 3) == 0 end with the sum of column lane >> 2 in p[0].
template <int CNT, int BIT>
__device__ __forceinline__ void colsum_halve(double (&p)[16], int lane) {
  const bool upper = (lane & BIT) != 0;
#pragma unroll
  for (int m = 0; m < CNT; ++m) {
    const double send = upper ? p[m] : p[m + CNT];
    const double keep = upper ? p[m + CNT] : p[m];
    p[m] = keep + __shfl_xor(send, BIT, 64);
  }
}
__device__ __forceinline__ void wave_colsum16(double (&p)[16], int lane) {
  colsum_halve<8, 32>(p, lane);
  colsum_halve<4, 16>(p, lane);
  colsum_halve<2, 8>(p, lane);
  colsum_halve<1, 4>(p, lane);
  p[0] += __shfl_xor(p[0], 2, 64);
  p[0] += __shfl_xor(p[0], 1, 64);
}

// Backward: x_k = Dinv_k^T (y_k - sum_{j>k} L_jk^T x_j).  Lanes run down the
// rows of each L_jk block (coalesced) and keep per-column partial sums,
// summed over the wave's lanes once, at the end (wave_colsum16).  Waits as
// in the forward pass.
template <int NR>
__global__ __launch_bounds__(256) void potrs_bwd_kernel(const double* __restrict__ L, int ld,
                                                        const double* __restrict__ dinv, double* X, int ldx, int nb,
                                                        int* sync, int gen) {
  __shared__ int k_s;
  __shared__ double xv[NR][64];
  __shared__ double cs[NR][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int* err = sync + 2;
  int* flags = sync + 3 + nb;
  if (tid == 0) {
    const int t = atomicAdd(&sync[1], 1);
    k_s = nb - 1 - t;
    if (t == nb - 1) __hip_atomic_store(&sync[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int k = k_s;
  // Dinv_k and y_k depend on no other workgroup: load them now, off the chain
  const double* D = dinv + (size_t)k * 4096;
  double dv[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) dv[c] = D[lane + 64 * (16 * wv + c)];
  const double yk = wv < NR ? X[(size_t)wv * ldx + (size_t)64 * k + lane] : 0.0;
  double p[NR][16];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < 16; ++c) p[r][c] = 0.0;
  for (int j = nb - 1; j > k; --j) {
    const double* Lb = L + (size_t)64 * j + lane + (size_t)(64 * k + 16 * wv) * ld;
    double lv[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) lv[c] = Lb[(size_t)c * ld];
    if (lane == 0) wait_flag(&flags[j], gen, err);
    __builtin_amdgcn_wave_barrier();
    wt_seen();
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double xr = wt_load(X + (size_t)r * ldx + (size_t)64 * j + lane);
#pragma unroll
      for (int c = 0; c < 16; ++c) p[r][c] += lv[c] * xr;
    }
  }
  // t = y_k - column sums (columns 16 wv + c of the wave)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    wave_colsum16(p[r], lane);
    if ((lane & 3) == 0) cs[r][16 * wv + (lane >> 2)] = p[r][0];
  }
  __syncthreads();
  if (wv < NR) xv[wv][lane] = yk - cs[wv][lane];
  __syncthreads();
  // x_k = Dinv_k^T t: the same column sums, of Dinv_k's columns times t
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const double t = xv[r][lane];
#pragma unroll
    for (int c = 0; c < 16; ++c) p[r][c] = dv[c] * t;
    wave_colsum16(p[r], lane);
    if ((lane & 3) == 0) wt_store(X + (size_t)r * ldx + (size_t)64 * k + 16 * wv + (lane >> 2), p[r][0]);
  }
  wt_stores_done();
  __syncthreads();
  if (tid == 0) wt_publish(&flags[k], gen);
}

// [0] forward ticket, [1] backward ticket, [2] error word, then the forward
// and the backward flags (nb each, holding the generation that set them).
// The tickets are cleared by each launch's last workgroup.
static int sync_reserve(DenseWork& w, int nb, hipStream_t s) {
  const size_t need = 3 + 2 * (size_t)nb;
  if (need > w.sync_cap) {
    if (w.d_sync) RSD_HIP(hipFree(w.d_sync));
    RSD_HIP(hipMalloc(&w.d_sync, need * sizeof(int)));
    RSD_HIP(hipMemsetAsync(w.d_sync, 0, need * sizeof(int), s));
    w.sync_cap = need;
    w.gen = 0;
  }
  return RSD_OK;
}

int potrs_lower(DenseWork& w, const double* A, int n, int ld, double* X, int ldx, int nrhs, hipStream_t s) {
  if (n % 64 || nrhs < 1 || nrhs > 2) {
    set_error("potrs_lower: n must be a multiple of 64 and nrhs 1 or 2");
    return RSD_ERR_ARG;
  }
  const int nb = n / 64;
  const int rc = sync_reserve(w, nb, s);
  if (rc) return rc;
  // flags hold the generation that set them; a new generation needs no clearing
  if (++w.gen == 0x7fffffff) w.gen = 1;
  hipEvent_t pe = prof_begin(kProfTrsv, 2.0 * n * n * nrhs / 2.0, 8.0 * n * n / 2.0, s);
  if (nrhs == 1)
    hipLaunchKernelGGL(potrs_fwd_kernel<1>, dim3(nb), dim3(256), 0, s, A, ld, (const double*)w.dinv, X, ldx, nb,
                       w.d_sync, w.gen);
  else
    hipLaunchKernelGGL(potrs_fwd_kernel<2>, dim3(nb), dim3(256), 0, s, A, ld, (const double*)w.dinv, X, ldx, nb,
                       w.d_sync, w.gen);
  prof_end(pe, s);
  pe = prof_begin(kProfTrsv, 2.0 * n * n * nrhs / 2.0, 8.0 * n * n / 2.0, s);
  if (nrhs == 1)
    hipLaunchKernelGGL(potrs_bwd_kernel<1>, dim3(nb), dim3(256), 0, s, A, ld, (const double*)w.dinv, X, ldx, nb,
                       w.d_sync, w.gen);
  else
    hipLaunchKernelGGL(potrs_bwd_kernel<2>, dim3(nb), dim3(256), 0, s, A, ld, (const double*)w.dinv, X, ldx, nb,
                       w.d_sync, w.gen);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int trsv_lower(DenseWork& w, const double* L, int n, int ld, const double* dinv, double* x, bool trans,
               hipStream_t s) {
  if (n % 64) {
    set_error("trsv_lower: n must be a multiple of 64");
    return RSD_ERR_ARG;
  }
  const int nb = n / 64;
  const int rc = sync_reserve(w, nb, s);
  if (rc) return rc;
  if (++w.gen == 0x7fffffff) w.gen = 1;
  hipEvent_t pe = prof_begin(kProfTrsv, (double)n * n, 8.0 * n * n / 2.0, s);
  if (trans)
    hipLaunchKernelGGL(potrs_bwd_kernel<1>, dim3(nb), dim3(256), 0, s, L, ld, dinv, x, n, nb, w.d_sync, w.gen);
  else
    hipLaunchKernelGGL(potrs_fwd_kernel<1>, dim3(nb), dim3(256), 0, s, L, ld, dinv, x, n, nb, w.d_sync, w.gen);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// G' = [C B^T; B A] (lower triangle) from G = [A B; B^T C]: C (n5 x n5) padded to
// n5p with identity, then A (n6 x n6) padded to n6p; B below C
// The padding of the reordered grand mobility [C B^T; B A] (order n5p + n6p):
// rows and columns n5..n5p-1 (C's) and n5p+n6..n5p+n6p-1 (A's) become the
// identity's.  One thread per (padding index, column).
__global__ void grand_padding_kernel(double* __restrict__ Gp, int ld, int n5, int n5p, int n6, int n6p) {
  const int np = n5p + n6p, pc = n5p - n5;
  const int k = blockIdx.y;  // the k-th padding index
  const int p = k < pc ? n5 + k : n5p + n6 + (k - pc);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < np; c += gridDim.x * blockDim.x) {
    const double v = c == p ? 1.0 : 0.0;
    Gp[(size_t)p + (size_t)c * ld] = v;
    Gp[(size_t)c + (size_t)p * ld] = v;
  }
}

int grand_padding(double* Gp, int ld, int n5, int n5p, int n6, int n6p, hipStream_t s) {
  const int npad = (n5p - n5) + (n6p - n6);
  if (npad == 0) return RSD_OK;
  hipLaunchKernelGGL(grand_padding_kernel, dim3((n5p + n6p + 255) / 256, npad), dim3(256), 0, s, Gp, ld, n5, n5p, n6,
                     n6p);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

const int* dense_solve_error_word(const DenseWork& w) { return w.d_sync ? w.d_sync + 2 : nullptr; }
const int* dense_factor_error_word(const DenseWork& w) { return w.d_dag ? w.d_dag + 1 : nullptr; }

int dense_factor_wait_record(const DenseWork& w, int rec[4]) {
  for (int k = 0; k < 4; ++k) rec[k] = 0;
  if (!w.d_rec) return RSD_OK;
  RSD_HIP(hipMemcpy(rec, w.d_rec, 4 * sizeof(int), hipMemcpyDeviceToHost));
  return RSD_OK;
}

int dense_wait_reset(DenseWork& w, hipStream_t s) {
  if (!w.d_dag || !w.d_rec) return RSD_OK;
  hipLaunchKernelGGL(dag_wait_reset_kernel, dim3(1), dim3(64), 0, s, w.d_dag + 1, w.d_rec);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int dense_sync_error(DenseWork& w, hipStream_t s, int* err) {
  int e[2] = {0, 0};
  if (w.d_sync) RSD_HIP(hipMemcpyAsync(&e[0], w.d_sync + 2, sizeof(int), hipMemcpyDeviceToHost, s));
  if (w.d_dag) RSD_HIP(hipMemcpyAsync(&e[1], w.d_dag + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipStreamSynchronize(s));
  *err = e[0] | e[1];
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

int trtri_lower(DenseWork& w, double* A, int n, int ld, hipStream_t s, const double* dinv) {
  const int nt = n / 64;
  hipLaunchKernelGGL(put_dinv_kernel, dim3(nt), dim3(256), 0, s, A, ld, dinv ? dinv : (const double*)w.dinv);
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

// FP32 copies for the CG preconditioner: the lower triangle of an FP64 factor
// (diagonal tiles excepted: they get the FP32 diagonal-tile inverses)
__global__ void lower_to_f32_kernel(const double* __restrict__ L, int ldl, float* __restrict__ F, int ldf, int n) {
  const int r = blockIdx.x * 64 + (threadIdx.x & 63), c = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r < n && c < n && (r >> 6) > (c >> 6)) F[(size_t)r + (size_t)c * ldf] = (float)L[(size_t)r + (size_t)c * ldl];
}

__global__ __launch_bounds__(256) void put_dinv32_kernel(float* A, int ld, const double* dinv) {
  const int k = blockIdx.x;
  for (int idx = threadIdx.x; idx < 4096; idx += 256) {
    const int r = idx & 63, c = idx >> 6;
    A[(size_t)(64 * k + r) + (size_t)(64 * k + c) * ld] = (float)dinv[(size_t)k * 4096 + r + (size_t)c * 64];
  }
}

int factor_to_f32(DenseWork& w, const double* L, int n, int ld, float* F, int ldf, hipStream_t s) {
  hipLaunchKernelGGL(lower_to_f32_kernel, dim3((n + 63) / 64, (n + 3) / 4), dim3(256), 0, s, L, ld, F, ldf, n);
  hipLaunchKernelGGL(put_dinv32_kernel, dim3(n / 64), dim3(256), 0, s, F, ldf, (const double*)w.dinv);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// trtri_lower's recursion in FP32, on a matrix whose diagonal tiles already
// hold their inverses (factor_to_f32); T is n x n FP32 scratch
int trtri_lower32(DenseWork& w, float* A, int n, int ld, float* T, hipStream_t s) {
  const int nt = n / 64;
  std::vector<std::vector<Node>> lv;
  build_levels(0, nt, 0, lv);
  for (int d = (int)lv.size() - 1; d >= 0; --d) {
    std::vector<Gemm32Desc> g1, g2;
    for (const Node& nd : lv[d]) {
      Gemm32Desc g{};
      g.A = A + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * ld;
      g.B = A + (size_t)(64 * nd.a) + (size_t)(64 * nd.a) * ld;
      g.C = T + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * n;
      g.lda = ld;
      g.ldb = ld;
      g.ldc = n;
      g.mt = nd.b - nd.m;
      g.nt = nd.m - nd.a;
      g.kt = nd.m - nd.a;
      g.alpha = 1.f;
      g.beta = 0.f;
      g1.push_back(g);
      Gemm32Desc h{};
      h.A = A + (size_t)(64 * nd.m) + (size_t)(64 * nd.m) * ld;
      h.B = T + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * n;
      h.C = A + (size_t)(64 * nd.m) + (size_t)(64 * nd.a) * ld;
      h.lda = ld;
      h.ldb = n;
      h.ldc = ld;
      h.mt = nd.b - nd.m;
      h.nt = nd.m - nd.a;
      h.kt = nd.b - nd.m;
      h.alpha = -1.f;
      h.beta = 0.f;
      g2.push_back(h);
    }
    RSD_TRY(gemm32_batched(w, g1, false, kKloCol, s));
    RSD_TRY(gemm32_batched(w, g2, false, kKhiRow, s));
  }
  return RSD_OK;
}

// the upper triangle of an n x n FP32 matrix from its lower one, through LDS
// so that both sides are read and written column-contiguously
__global__ __launch_bounds__(256) void sym_fill_upper_f32_kernel(float* P, int n) {
  const int R = blockIdx.x, C = blockIdx.y;  // destination block of 32 x 32
  if (R >= C) return;
  __shared__ float t[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) t[j][tx] = P[(size_t)(C * 32 + tx) + (size_t)(R * 32 + j) * n];
  __syncthreads();
  for (int j = ty; j < 32; j += 8) P[(size_t)(R * 32 + tx) + (size_t)(C * 32 + j) * n] = t[tx][j];
}

int sym_fill_upper_f32(float* P, int n, hipStream_t s) {
  hipLaunchKernelGGL(sym_fill_upper_f32_kernel, dim3(n / 32, n / 32), dim3(256), 0, s, P, n);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// C = X^T X (lower tiles), X lower-triangular FP32
int lauum_lower32(DenseWork& w, const float* X, int ldx, float* C, int ldc, int n, hipStream_t s) {
  const int nt = n / 64;
  Gemm32Desc g{};
  g.A = X;
  g.B = X;
  g.C = C;
  g.lda = ldx;
  g.ldb = ldx;
  g.ldc = ldc;
  g.mt = g.nt = g.kt = nt;
  g.alpha = 1.f;
  g.beta = 0.f;
  return gemm32_batched(w, {g}, true, kLowerC | kKloRow, s);
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

__global__ void zero_upper_kernel(double* A, int n, int ld) {
  const int c = blockIdx.y * 16 + threadIdx.y;
  const int r = blockIdx.x * 16 + threadIdx.x;
  if (r < n && c < n && r < c) A[(size_t)r + (size_t)c * ld] = 0.0;
}

int zero_strict_upper(double* A, int n, int ld, hipStream_t s) {
  dim3 b(16, 16), g((n + 15) / 16, (n + 15) / 16);
  hipLaunchKernelGGL(zero_upper_kernel, g, b, 0, s, A, n, ld);
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

__global__ void vector_sum_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ y,
                                  int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] + b[i];
}

int vector_sum(const double* a, const double* b, double* y, int n, hipStream_t s) {
  if (n <= 0) return RSD_OK;
  hipLaunchKernelGGL(vector_sum_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, b, y, n);
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

// 64 rows per workgroup: wave w sums chunks w, w+4, ... in order, then the
// four wave sums combine in a fixed order (bitwise repeatable)
__global__ __launch_bounds__(256) void gemv_n_reduce_kernel(int m, int nchunk, const double* __restrict__ part,
                                                            double alpha, double beta, double* __restrict__ y) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (r < m) {
#pragma unroll 4
    for (int c = wv; c < nchunk; c += 4) acc += part[(size_t)c * m + r];
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && r < m) {
    const double tot = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    y[r] = alpha * tot + (beta != 0.0 ? beta * y[r] : 0.0);
  }
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
    hipLaunchKernelGGL(gemv_n_reduce_kernel, dim3((m + 63) / 64), dim3(256), 0, s, m, nchunk,
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
  if (w.d_sync) (void)hipFree(w.d_sync);
  if (w.d_dag) (void)hipFree(w.d_dag);
  if (w.d_rec) (void)hipFree(w.d_rec);
  if (w.d_part) (void)hipFree(w.d_part);
  w = DenseWork{};
}

}  // namespace rsd
