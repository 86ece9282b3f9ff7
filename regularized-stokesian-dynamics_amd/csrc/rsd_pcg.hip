// rsd_pcg.hip -- preconditioned conjugate gradients for the collision
// sub-step solves.
//
// Between the NT_DEM collision sub-steps of one step (stepper.f90:95-143)
// only the lubrication part of RPP = inv(M)_PP + R_lub moves, by the
// sub-step displacement.  The first sub-step factors RPP_1 exactly as the
// direct path does; its inverse, rounded to FP32, then preconditions CG on
// RPP_k for the later sub-steps, started from the last two sub-steps'
// solutions extrapolated and run in FP64 to a relative residual of 1e-14.  The
// iteration count is about 15-30 at N=512, phi=0.4 (two matrix-vector
// products each, both HBM-bound), against a 6N factorisation per sub-step.
//
// One iteration is two launches, each one matrix-vector product with the
// vector algebra of CG fused around it:
//   direction:  p_k = z_k + beta p_{k-1};  q = A p_k;        partial p.q
//   update:     alpha = rho_k / p.q;  x += alpha p_k;  r_{k+1} = r_k - alpha q;
//               z_{k+1} = P r_{k+1};  partials r.z and r.r
// A and P are symmetric, so y = A v is formed column by column (y_c = A_:c . v):
// one wavefront per column, lanes down contiguous rows, 16-byte loads, the
// first batch of them issued before the launch's scalars are known.  The
// operand vector is formed once per workgroup in LDS.  Every workgroup
// reduces the previous launch's per-workgroup partial sums itself, in one
// fixed order, so all of them take the same scalars and the same decisions
// (converged, breakdown) without atomics.
#include "rsd_internal.h"
#include "rsd_pcg.h"
#include "rsd_prof.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace rsd {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;  // one wavefront per column: kWaves columns per workgroup
constexpr int kStage = 8192;           // rows of the operand vector held in LDS at once
constexpr int kBatch = 8;              // 16-byte matrix loads in flight per lane (per buffer)
constexpr int kPhaseBatch = 12;        // the same for the one-workgroup-per-CU phases (a whole column
                                       // per batch at N=512 in FP32, half a column in FP64)

enum Mode : int { kResidual = 0, kPrecond0 = 1, kDirection = 2, kUpdate = 3 };

// partial-sum slots, each nblk doubles
enum Slot : int { kSigma = 0, kRho0 = 1, kRho1 = 2, kRr0 = 3, kRr1 = 4, kBb = 5, kSlots = 6 };

struct PcgArgs {
  const double* A;  // FP64 operator, column-major, leading dim lda
  const float* P;   // FP32 preconditioner, leading dim ldp
  int lda, ldp;
  double* x;
  const double* b;
  double* r[2];
  double* z;
  double* p[2];
  double* q;
  double* part;
  int* state;       // [0] status (PcgStatus), [1] iterations when it stopped
  double tol2;
  int n, nblk, k;
};

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double2 type;
  static constexpr int kElems = 2;
};
template <>
struct Vec16<float> {
  typedef float4 type;
  static constexpr int kElems = 4;
};

__device__ inline double dot16(const double2& a, const double* v) { return fma(a.x, v[0], a.y * v[1]); }
__device__ inline double dot16(const float4& a, const double* v) {
  return fma((double)a.x, v[0], fma((double)a.y, v[1], fma((double)a.z, v[2], (double)a.w * v[3])));
}

__device__ inline double dir_elem(const PcgArgs& g, int i, double beta) {
  return g.k == 0 ? g.z[i] : fma(beta, g.p[(g.k + 1) & 1][i], g.z[i]);
}

__device__ inline double upd_elem(const PcgArgs& g, int i, double alpha) {
  return fma(-alpha, g.q[i], g.r[g.k & 1][i]);
}

// The operand of this launch's matrix-vector product is f(a_i, b_i) for two
// vectors a, b and the launch's scalars.
template <int MODE>
__device__ inline const double* operand_a(const PcgArgs& g) {
  return MODE == kResidual ? g.x : MODE == kPrecond0 ? g.r[0] : MODE == kDirection ? g.z : g.r[g.k & 1];
}
template <int MODE>
__device__ inline const double* operand_b(const PcgArgs& g) {
  return MODE == kDirection ? g.p[(g.k + 1) & 1] : MODE == kUpdate ? g.q : g.x;
}
template <int MODE>
__device__ inline double operand_f(const PcgArgs& g, double a, double b, double alpha, double beta) {
  if (MODE == kDirection) return g.k == 0 ? a : fma(beta, b, a);
  if (MODE == kUpdate) return fma(-alpha, b, a);
  return a;
}

constexpr int kGroup = 16;  // operand elements per thread held in registers while the first matrix batch flies

template <int MODE, typename T>
__global__ __launch_bounds__(kThreads) void pcg_kernel(PcgArgs g) {
  extern __shared__ double vs[];  // min(n, kStage) doubles
  __shared__ double red[kWaves][4];
  __shared__ double colp[kWaves][2];
  typedef typename Vec16<T>::type V;
  constexpr int E = Vec16<T>::kElems;
  constexpr int K = MODE == kDirection ? 4 : MODE == kUpdate ? 2 : 1;  // scalars from partial sums
  const int n = g.n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = blockIdx.x * kWaves + w;  // the column (= output row) of this wavefront
  const T* Mbase;
  int ld;
  if constexpr (std::is_same<T, double>::value) {
    Mbase = g.A;
    ld = g.lda;
  } else {
    Mbase = g.P;
    ld = g.ldp;
  }
  const V* cv = (const V*)(Mbase + (size_t)c * ld);
  int slot[K];
  if (MODE == kDirection) {
    slot[0] = kRr0 + (g.k & 1);
    slot[1] = kBb;
    slot[2] = kRho0 + (g.k & 1);
    slot[3] = kRho0 + ((g.k + 1) & 1);
  } else if (MODE == kUpdate) {
    slot[0] = kSigma;
    slot[1] = kRho0 + (g.k & 1);
  } else {
    slot[0] = kSigma;
  }
  // 1. Loads, in the order their values are needed (the memory counter is
  //    in order): status, the previous launch's partials, the operand's
  //    inputs, then the first batch of matrix columns.  Indices are clamped
  //    instead of branched around, and clamped values are masked.
  const int st = (MODE == kDirection || MODE == kUpdate) ? g.state[0] : (int)kPcgRunning;
  double pt[K][4];
  if (MODE == kDirection || MODE == kUpdate) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = min(j * kThreads + tid, g.nblk - 1);
#pragma unroll
      for (int k = 0; k < K; ++k) pt[k][j] = g.part[(size_t)slot[k] * g.nblk + i];
    }
  }
  const int rn0 = min(n, kStage);
  const double* oa = operand_a<MODE>(g);
  const double* ob = operand_b<MODE>(g);
  constexpr bool kTwo = MODE == kDirection || MODE == kUpdate;
  double va[kGroup], vb[kGroup];
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    const int i = min(j * kThreads + tid, rn0 - 1);
    va[j] = oa[i];
    vb[j] = kTwo ? ob[i] : 0.0;
  }
  if (st != kPcgRunning) return;
  // 16-byte vectors in the chunk, and this lane's share (vectors lane + 64 i);
  // loads past the end are clamped to the chunk's last vector and not used
  int nvc = rn0 / E;
  int nvl = (nvc - lane + 63) / 64;
  V cur[kBatch];
#pragma unroll
  for (int u = 0; u < kBatch; ++u) cur[u] = cv[min(lane + 64 * u, nvc - 1)];
  // 2. the scalars: a strided sum in index order, a fixed shuffle tree and the
  //    wave sums in order, the same in every thread of every workgroup
  double alpha = 0.0, beta = 0.0;
  if (MODE == kDirection || MODE == kUpdate) {
    double sum[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      sum[k] = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j * kThreads + tid < g.nblk) sum[k] += pt[k][j];
    }
    for (int i0 = 4 * kThreads; i0 < g.nblk; i0 += kThreads) {  // more than 4 * kThreads workgroups
      const int i = i0 + tid;
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (i < g.nblk) sum[k] += g.part[(size_t)slot[k] * g.nblk + i];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      for (int off = 32; off > 0; off >>= 1) sum[k] += __shfl_xor(sum[k], off, 64);
      if (lane == 0) red[w][k] = sum[k];
    }
    __syncthreads();
    double tot[K];
#pragma unroll
    for (int k = 0; k < K; ++k) tot[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
    if (MODE == kDirection) {
      if (tot[0] <= g.tol2 * tot[1]) {
        if (blockIdx.x == 0 && tid == 0) {
          g.state[0] = kPcgConverged;
          g.state[1] = g.k;
        }
        return;
      }
      if (g.k > 0) beta = tot[2] / tot[3];
    } else {
      const double sigma = tot[0], rho = tot[1];
      if (!(sigma > 0.0) || !(rho >= 0.0)) {  // not SPD (or NaN): the caller falls back to the factorisation
        if (blockIdx.x == 0 && tid == 0) {
          g.state[0] = kPcgBreakdown;
          g.state[1] = g.k;
        }
        return;
      }
      alpha = rho / sigma;
    }
  }
  // 3. chunks of the operand through LDS; matrix batches double-buffered
  double acc = 0.0;
  for (int r0 = 0; r0 < n; r0 += kStage) {
    const int rn = min(kStage, n - r0);
    if (r0 > 0) {
      nvc = rn / E;
      nvl = (nvc - lane + 63) / 64;
      cv += kStage / E;
#pragma unroll
      for (int u = 0; u < kBatch; ++u) cur[u] = cv[min(lane + 64 * u, nvc - 1)];
      __syncthreads();  // the previous chunk's readers are done
    }
    for (int i0 = 0; i0 < rn; i0 += kGroup * kThreads) {
      if (r0 > 0 || i0 > 0) {
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
          const int i = min(i0 + j * kThreads + tid, rn - 1);
          va[j] = oa[r0 + i];
          vb[j] = kTwo ? ob[r0 + i] : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < kGroup; ++j) {
        const int i = i0 + j * kThreads + tid;
        if (i < rn) vs[i] = operand_f<MODE>(g, va[j], vb[j], alpha, beta);
      }
    }
    __syncthreads();
    for (int i0 = 0; i0 < nvl; i0 += kBatch) {
      V nxt[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) nxt[u] = cv[min(lane + 64 * (i0 + kBatch + u), nvc - 1)];
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (i0 + u < nvl) acc += dot16(cur[u], vs + (size_t)(lane + 64 * (i0 + u)) * E);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) cur[u] = nxt[u];
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  // 4. lane 0 finishes column c; thread 0 writes the workgroup's partials
  if (lane == 0) {
    const double y = acc;
    double pa = 0.0, pb = 0.0;
    if (MODE == kResidual) {
      const double rc = g.b[c] - y;
      g.r[0][c] = rc;
      pa = rc * rc;
      pb = g.b[c] * g.b[c];
    } else if (MODE == kPrecond0) {
      g.z[c] = y;
      pa = g.r[0][c] * y;
    } else if (MODE == kDirection) {
      const double pc = dir_elem(g, c, beta);
      g.p[g.k & 1][c] = pc;
      g.q[c] = y;
      pa = pc * y;
    } else {
      const double rc = upd_elem(g, c, alpha);
      g.r[(g.k + 1) & 1][c] = rc;
      g.x[c] = fma(alpha, g.p[g.k & 1][c], g.x[c]);
      g.z[c] = y;
      pa = rc * y;
      pb = rc * rc;
    }
    colp[w][0] = pa;
    colp[w][1] = pb;
  }
  __syncthreads();
  if (tid == 0) {
    const double sa = (colp[0][0] + colp[1][0]) + (colp[2][0] + colp[3][0]);
    const double sb = (colp[0][1] + colp[1][1]) + (colp[2][1] + colp[3][1]);
    const size_t nb = g.nblk, bi = blockIdx.x;
    if (MODE == kResidual) {
      g.part[kRr0 * nb + bi] = sa;
      g.part[kBb * nb + bi] = sb;
      if (bi == 0) {
        g.state[0] = kPcgRunning;
        g.state[1] = 0;
      }
    } else if (MODE == kPrecond0) {
      g.part[kRho0 * nb + bi] = sa;
    } else if (MODE == kDirection) {
      g.part[kSigma * nb + bi] = sa;
    } else {
      const int s1 = (g.k + 1) & 1;
      g.part[(kRho0 + s1) * nb + bi] = sa;
      g.part[(kRr0 + s1) * nb + bi] = sb;
    }
  }
}

// ------------------------------------------------- one-launch CG
// The whole iteration in one cooperative launch: one workgroup per CU, the
// phases above separated by grid barriers (two per iteration), every
// workgroup taking the same decisions from the same partial sums, so all of
// them leave the loop together.  A wavefront owns columns gw, gw + nw, ...
// of the matrix; the first batch of its next phase's matrix loads is in
// flight while it waits at the barrier.
constexpr int kPersistMaxN = 16384;  // operand vector held in LDS whole (128 KB)

struct CgArgs {
  const double* A;
  const float* P;
  int lda, ldp;
  double* x;
  const double* b;
  double* r[2];
  double* z;
  double* p[2];
  double* q;
  double* part;     // kSlots x G partials
  int* state;       // [0] PcgStatus, [1] iterations
  unsigned* bar;    // barrier arrivals; zero at launch
  int* err;         // set when a barrier wait ran out
  double tol2;
  int n, G, max_iter;
};

constexpr long kBarrierSpin = 1L << 22;

// Grid barrier number `nbar`: every workgroup's writes before it are visible
// to every workgroup after it.  Two levels: workgroups arrive on one of 16
// group counters (bar[64 (1 + grp)]), the last of a group on the top counter
// bar[0], which everyone polls.  Bounded: a wait that runs out sets *err, and
// once *err is set every wait returns at once, so the launch always ends.
constexpr int kBarGroups = 16;
constexpr int kBarErr = 64 * (kBarGroups + 1);  // the error word's index in bar
constexpr int kBarWords = kBarErr + 64;
__device__ void grid_sync(const CgArgs& g, unsigned nbar) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const int grp = blockIdx.x % kBarGroups;
    const unsigned members = (unsigned)((g.G - grp + kBarGroups - 1) / kBarGroups);
    const unsigned old =
        __hip_atomic_fetch_add(g.bar + 64 * (1 + grp), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == nbar * members) __hip_atomic_fetch_add(g.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = nbar * (unsigned)min(g.G, kBarGroups);
    long it = 0;
    while (__hip_atomic_load(g.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if ((++it & 255) == 0 && __hip_atomic_load(g.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (it >= kBarrierSpin) {
        atomicOr(g.err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

template <int MODE>
__device__ inline const double* cg_operand_a(const CgArgs& g, int k) {
  return MODE == kResidual ? g.x : MODE == kPrecond0 ? g.r[0] : MODE == kDirection ? g.z : g.r[k & 1];
}
template <int MODE>
__device__ inline const double* cg_operand_b(const CgArgs& g, int k) {
  return MODE == kDirection ? g.p[(k + 1) & 1] : MODE == kUpdate ? g.q : g.x;
}
template <int MODE>
__device__ inline double cg_operand_f(int k, double a, double b, double alpha, double beta) {
  if (MODE == kDirection) return k == 0 ? a : fma(beta, b, a);
  if (MODE == kUpdate) return fma(-alpha, b, a);
  return a;
}

// One phase; returns true when the solve ends here (all workgroups alike).
// PHASED: the phase is a launch of its own (the kernel boundary is the
// barrier).  Launches past the end of a solve must do nothing, and they know
// it from the partials they read anyway: a phase that ends the solve, or
// finds it ended, writes NaN into the partials the next phase reads (direction
// -> sigma; update -> rho and r.r).  A NaN from the data itself reads the
// same way and leaves the status "running", which the caller treats as not
// converged.
template <int MODE, typename T, bool PHASED>
__device__ bool cg_phase(const CgArgs& g, int k, unsigned& nbar, double* vs, double (*red)[4],
                         double (*colp)[2]) {
  typedef typename Vec16<T>::type V;
  constexpr int E = Vec16<T>::kElems;
  constexpr int K = MODE == kDirection ? 4 : MODE == kUpdate ? 2 : 1;
  const int n = g.n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nw = g.G * kWaves, gw = blockIdx.x * kWaves + w;
  const int ncw = gw < n ? (n - 1 - gw) / nw + 1 : 0;  // this wavefront's columns
  const int nvc = n / E;                                // 16-byte vectors per column
  const int nvm = (nvc + 63) / 64;                      // per lane (the last ones masked)
  const T* M;
  int ld;
  if constexpr (std::is_same<T, double>::value) {
    M = g.A;
    ld = g.lda;
  } else {
    M = g.P;
    ld = g.ldp;
  }
  // address of this lane's vector i of its column j, clamped into the matrix
  auto vaddr = [&](int j, int i) -> const V* {
    const int col = min(gw + min(j, max(ncw - 1, 0)) * nw, n - 1);
    return (const V*)(M + (size_t)col * ld) + min(lane + 64 * i, nvc - 1);
  };
  // what lane j needs to finish this wavefront's column j: loaded now (its own
  // columns, written by this same lane in earlier phases, or b)
  const int cj = gw + lane * nw;
  const bool ownj = lane < ncw;
  double e0 = 0.0, e1 = 0.0;
  if (MODE == kResidual && ownj) e0 = g.b[cj];
  if (MODE == kUpdate && ownj) {
    e0 = g.x[cj];
    e1 = g.p[k & 1][cj];
  }
  // Loads go out in the order their values are needed (the memory counter
  // is in order): the previous phase's partials and the operand's inputs,
  // then the first batch of matrix columns.  In one launch the matrix batch
  // is issued before the barrier instead, to fly while the barrier waits.
  V cur[kPhaseBatch];
  if (!PHASED && MODE != kResidual) {
#pragma unroll
    for (int u = 0; u < kPhaseBatch; ++u) cur[u] = *vaddr(0, u);
    grid_sync(g, ++nbar);
  }
  int slot[K];
  if constexpr (MODE == kDirection) {
    slot[0] = kRr0 + (k & 1);
    slot[1] = kBb;
    slot[2] = kRho0 + (k & 1);
    slot[3] = kRho0 + ((k + 1) & 1);
  } else if constexpr (MODE == kUpdate) {
    slot[0] = kSigma;
    slot[1] = kRho0 + (k & 1);
  } else {
    slot[0] = kSigma;
  }
  constexpr bool kScalars = MODE == kDirection || MODE == kUpdate;
  double sum[K];
#pragma unroll
  for (int q = 0; q < K; ++q) sum[q] = kScalars ? g.part[(size_t)slot[q] * g.G + min(tid, g.G - 1)] : 0.0;
  const double* oa = cg_operand_a<MODE>(g, k);
  const double* ob = cg_operand_b<MODE>(g, k);
  constexpr bool kTwo = MODE == kDirection || MODE == kUpdate;
  double va[kGroup], vb[kGroup];
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    const int i = min(j * kThreads + tid, n - 1);
    va[j] = oa[i];
    vb[j] = kTwo ? ob[i] : 0.0;
  }
  if (PHASED || MODE == kResidual) {
#pragma unroll
    for (int u = 0; u < kPhaseBatch; ++u) cur[u] = *vaddr(0, u);
  }
  // scalars of this phase from the previous phase's partials
  double alpha = 0.0, beta = 0.0;
  if (kScalars) {
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (tid >= g.G) sum[q] = 0.0;
    for (int i = kThreads + tid; i < g.G; i += kThreads) {  // more workgroups than threads
#pragma unroll
      for (int q = 0; q < K; ++q) sum[q] += g.part[(size_t)slot[q] * g.G + i];
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      for (int off = 32; off > 0; off >>= 1) sum[q] += __shfl_xor(sum[q], off, 64);
      if (lane == 0) red[w][q] = sum[q];
    }
    __syncthreads();
    double tot[K];
#pragma unroll
    for (int q = 0; q < K; ++q) tot[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    const double nan = __builtin_nan("");
    if (MODE == kDirection) {
      if (PHASED && tot[0] != tot[0]) {  // ended earlier
        if (tid == 0) g.part[(size_t)kSigma * g.G + blockIdx.x] = nan;
        return true;
      }
      const bool conv = tot[0] <= g.tol2 * tot[1];
      if (conv || k >= g.max_iter) {
        if (blockIdx.x == 0 && tid == 0) {
          g.state[0] = conv ? kPcgConverged : kPcgNotConverged;
          g.state[1] = k;
        }
        if (PHASED && tid == 0) g.part[(size_t)kSigma * g.G + blockIdx.x] = nan;
        return true;
      }
      if (k > 0) beta = tot[2] / tot[3];
    } else {
      const double sigma = tot[0], rho = tot[1];
      if (!(sigma > 0.0) || !(rho >= 0.0)) {
        if (blockIdx.x == 0 && tid == 0 && sigma == sigma && rho == rho) {  // not SPD
          g.state[0] = kPcgBreakdown;
          g.state[1] = k;
        }
        if (PHASED && tid == 0) {
          const int s1 = (k + 1) & 1;
          g.part[(size_t)(kRho0 + s1) * g.G + blockIdx.x] = nan;
          g.part[(size_t)(kRr0 + s1) * g.G + blockIdx.x] = nan;
        }
        return true;
      }
      alpha = rho / sigma;
    }
  }
  // the operand, whole, into LDS (the first kGroup * kThreads rows are in registers)
  for (int i0 = 0; i0 < n; i0 += kGroup * kThreads) {
    if (i0 > 0) {
#pragma unroll
      for (int j = 0; j < kGroup; ++j) {
        const int i = min(i0 + j * kThreads + tid, n - 1);
        va[j] = oa[i];
        vb[j] = kTwo ? ob[i] : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const int i = i0 + j * kThreads + tid;
      if (i < n) vs[i] = cg_operand_f<MODE>(k, va[j], vb[j], alpha, beta);
    }
  }
  __syncthreads();
  // this wavefront's columns, batches double-buffered across column ends;
  // lane j keeps column j's product
  double yj = 0.0;
  for (int j = 0; j < ncw; ++j) {
    double acc = 0.0;
    for (int i0 = 0; i0 < nvm; i0 += kPhaseBatch) {
      V nxt[kPhaseBatch];
      const bool same = i0 + kPhaseBatch < nvm;
#pragma unroll
      for (int u = 0; u < kPhaseBatch; ++u) nxt[u] = same ? *vaddr(j, i0 + kPhaseBatch + u) : *vaddr(j + 1, u);
#pragma unroll
      for (int u = 0; u < kPhaseBatch; ++u) {
        const int e = lane + 64 * (i0 + u);
        if (i0 + u < nvm && e < nvc) acc += dot16(cur[u], vs + (size_t)e * E);
      }
#pragma unroll
      for (int u = 0; u < kPhaseBatch; ++u) cur[u] = nxt[u];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == j) yj = acc;
  }
  // lanes j < ncw finish their columns together; the operand value of column
  // c is vs[c] (p_k in the direction phase, r_{k+1} in the update phase)
  double pa = 0.0, pb = 0.0;
  if (ownj) {
    const int c = cj;
    const double y = yj, v = vs[c];
    if (MODE == kResidual) {
      const double rc = e0 - y;
      g.r[0][c] = rc;
      pa = rc * rc;
      pb = e0 * e0;
    } else if (MODE == kPrecond0) {
      g.z[c] = y;
      pa = v * y;
    } else if (MODE == kDirection) {
      g.p[k & 1][c] = v;
      g.q[c] = y;
      pa = v * y;
    } else {
      g.r[(k + 1) & 1][c] = v;
      g.x[c] = fma(alpha, e1, e0);
      g.z[c] = y;
      pa = v * y;
      pb = v * v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    pa += __shfl_xor(pa, off, 64);
    pb += __shfl_xor(pb, off, 64);
  }
  if (lane == 0) {
    colp[w][0] = pa;
    colp[w][1] = pb;
  }
  __syncthreads();
  if (tid == 0) {
    const double sa = (colp[0][0] + colp[1][0]) + (colp[2][0] + colp[3][0]);
    const double sb = (colp[0][1] + colp[1][1]) + (colp[2][1] + colp[3][1]);
    const size_t G = g.G, bi = blockIdx.x;
    if (MODE == kResidual) {
      g.part[kRr0 * G + bi] = sa;
      g.part[kBb * G + bi] = sb;
      if (bi == 0) {
        g.state[0] = kPcgRunning;
        g.state[1] = 0;
      }
    } else if (MODE == kPrecond0) {
      g.part[kRho0 * G + bi] = sa;
    } else if (MODE == kDirection) {
      g.part[kSigma * G + bi] = sa;
    } else {
      const int s1 = (k + 1) & 1;
      g.part[(kRho0 + s1) * G + bi] = sa;
      g.part[(kRr0 + s1) * G + bi] = sb;
    }
  }
  return false;
}

__global__ __launch_bounds__(kThreads) void cg_persistent_kernel(CgArgs g) {
  extern __shared__ double vs[];  // n doubles
  __shared__ double red[kWaves][4];
  __shared__ double colp[kWaves][2];
  unsigned nbar = 0;
  cg_phase<kResidual, double, false>(g, 0, nbar, vs, red, colp);
  cg_phase<kPrecond0, float, false>(g, 0, nbar, vs, red, colp);
  for (int k = 0;; ++k) {
    if (cg_phase<kDirection, double, false>(g, k, nbar, vs, red, colp)) break;
    if (cg_phase<kUpdate, float, false>(g, k, nbar, vs, red, colp)) break;
  }
}

template <int MODE, typename T>
__global__ __launch_bounds__(kThreads) void cg_phase_kernel(CgArgs g, int k) {
  extern __shared__ double vs[];  // n doubles
  __shared__ double red[kWaves][4];
  __shared__ double colp[kWaves][2];
  unsigned nbar = 0;
  cg_phase<MODE, T, true>(g, k, nbar, vs, red, colp);
}

// x := 2 x - xh (linear extrapolation from the two previous sub-steps'
// solutions) and xh := the old x; with first, only xh := x.
__global__ void extrapolate_kernel(double* __restrict__ x, double* __restrict__ xh, int n, int first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i];
  if (!first) x[i] = 2.0 * a - xh[i];
  xh[i] = a;
}

// the barrier counters (bar[0] and the group counters bar[64 (1 + grp)]) start from zero
__global__ void zero_barrier_kernel(unsigned* bar) {
  if (threadIdx.x <= kBarGroups) bar[64 * threadIdx.x] = 0u;
}

// A matrix whose lower triangle is valid: its strict upper triangle filled
// from the lower one (in place; only lower entries are read), and a full
// symmetric FP32 copy.
template <int MODE>
int launch_mode(const PcgArgs& g, hipStream_t s) {
  // the operator A in FP64 for the residual and direction launches, the FP32 preconditioner for the others
  typedef typename std::conditional<MODE == kResidual || MODE == kDirection, double, float>::type T;
  const size_t lds = (size_t)std::min(g.n, kStage) * sizeof(double);
  hipLaunchKernelGGL((pcg_kernel<MODE, T>), dim3(g.nblk), dim3(kThreads), lds, s, g);
  return RSD_OK;
}

}  // namespace

int pcg_reserve(PcgWork& pw, int n) {
  if (n % 64) {
    set_error("pcg: order must be a multiple of 64");
    return RSD_ERR_ARG;
  }
  if (!pw.probed) {
    // one workgroup per CU for the one-launch solver, if the device runs
    // cooperative launches and a CU holds the workgroup
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    RSD_HIP(hipGetDevice(&dev));
    RSD_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    RSD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RSD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cg_persistent_kernel, kThreads,
                                                         kPersistMaxN * sizeof(double)));
    pw.persist_G = (coop && per_cu >= 1) ? cus : 0;
    pw.cus = cus;
    pw.probed = true;
  }
  if (n <= pw.n_cap) return RSD_OK;
  float* P = nullptr;
  double *vec = nullptr, *part = nullptr;
  RSD_HIP(hipMalloc(&P, (size_t)n * n * sizeof(float)));
  RSD_HIP(hipMalloc(&vec, (size_t)7 * n * sizeof(double)));
  // the phased solvers index the partials by slot * (workgroups): n / kWaves
  // of them in the general form, one per CU (persist_G, or the CU count when
  // cooperative launches are not available) in the other two
  const int nblk = std::max(n / kWaves, std::max(pw.persist_G, pw.cus));
  RSD_HIP(hipMalloc(&part, (size_t)kSlots * nblk * sizeof(double)));
  if (pw.P) (void)hipFree(pw.P);
  if (pw.vec) (void)hipFree(pw.vec);
  if (pw.part) (void)hipFree(pw.part);
  pw.P = P;
  pw.vec = vec;
  pw.part = part;
  pw.part_blk = nblk;
  if (!pw.state) {

    RSD_HIP(hipMalloc(&pw.bar, kBarWords * sizeof(unsigned)));
    RSD_HIP(hipMemset(pw.bar, 0, kBarWords * sizeof(unsigned)));
    // the status slots live in pinned host memory that the kernels write
    // directly: a status check is a wait, with no copy before it
    RSD_HIP(hipHostMalloc(&pw.h_state, 2 * kPcgMaxSlots * sizeof(int), hipHostMallocMapped));
    RSD_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&pw.state), pw.h_state, 0));
  }
  pw.n_cap = n;
  pw.n = 0;
  return RSD_OK;
}

void pcg_release(PcgWork& pw) {
  if (pw.P) (void)hipFree(pw.P);
  if (pw.vec) (void)hipFree(pw.vec);
  if (pw.part) (void)hipFree(pw.part);
  if (pw.bar) (void)hipFree(pw.bar);
  if (pw.h_state) (void)hipHostFree(pw.h_state);
  pw = PcgWork{};
}

int pcg_precond_from_factor(DenseWork& w, PcgWork& pw, const double* L, int n, int ld, hipStream_t s,
                            hipEvent_t copied) {
  RSD_TRY(pcg_reserve(pw, n));
  // inv(L L^T) = L^-T L^-1 in FP32: the factor's FP32 copy (its diagonal
  // tiles the potrf inverses) and the triangular inverse's scratch share w's
  // n x n FP64 scratch; the Gram product lands in P
  float* F = reinterpret_cast<float*>(w.tmp);
  float* T = F + (size_t)n * n;
  RSD_TRY(factor_to_f32(w, L, n, ld, F, n, s));
  if (copied) RSD_HIP(hipEventRecord(copied, s));
  RSD_TRY(trtri_lower32(w, F, n, n, T, s));
  RSD_TRY(lauum_lower32(w, F, n, pw.P, n, n, s));
  RSD_TRY(sym_fill_upper_f32(pw.P, n, s));
  pw.n = n;
  return RSD_OK;
}

int pcg_initial_guess(double* x, double* xh, int n, bool first, hipStream_t s) {
  hipLaunchKernelGGL(extrapolate_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, xh, n, first ? 1 : 0);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

bool pcg_one_launch(const PcgWork& pw, int n) { return pw.persist_G > 0 && n <= kPersistMaxN; }

// Waits until the solve of a status slot has ended (the kernels have written
// its status into pinned host memory) or the stream has drained, whichever
// comes first; *ended tells which.  The host thus enqueues the work after a
// solve as soon as it converges, while the launches enqueued past convergence
// (which return at once) still run, instead of letting the stream drain.
static int wait_solve_end(const PcgWork& pw, int slot, hipStream_t s, bool* ended) {
  const int* st = pw.h_state + 2 * slot;
  for (unsigned spin = 0;; ++spin) {
    if (__atomic_load_n(st, __ATOMIC_ACQUIRE) != kPcgRunning) {
      *ended = true;
      return RSD_OK;
    }
    if ((spin & 255) == 255) {
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess) {
        *ended = __atomic_load_n(st, __ATOMIC_ACQUIRE) != kPcgRunning;
        return RSD_OK;
      }
      if (q != hipErrorNotReady) RSD_HIP(q);
    }
    __builtin_ia32_pause();
  }
}

int pcg_solve(PcgWork& pw, const double* A, int n, int lda, const double* b, double* x, double tol, int max_iter,
              int slot, bool one_launch, int chunk_hint, hipStream_t s) {
  if (n != pw.n || !pw.P) {
    set_error("pcg: no preconditioner of this order");
    return RSD_ERR_ARG;
  }
  if (slot < 0 || slot >= kPcgMaxSlots) {
    set_error("pcg: status slot out of range");
    return RSD_ERR_ARG;
  }
  int* state = pw.state + 2 * slot;
  // The slot holds the status of an earlier solve until this one's first
  // launch runs: marked running here, so that the wait below cannot take it
  // for this solve's end (every launch that wrote it has completed, since the
  // slot is used once per step and steps end with the stream synchronised)
  __atomic_store_n(pw.h_state + 2 * slot, (int)kPcgRunning, __ATOMIC_RELEASE);
  hipEvent_t pe = prof_begin(kProfPcg, 0.0, 0.0, s);
  CgArgs cg{};
  cg.A = A;
  cg.lda = lda;
  cg.P = pw.P;
  cg.ldp = n;
  cg.x = x;
  cg.b = b;
  cg.r[0] = pw.vec;
  cg.r[1] = pw.vec + n;
  cg.z = pw.vec + 2 * (size_t)n;
  cg.p[0] = pw.vec + 3 * (size_t)n;
  cg.p[1] = pw.vec + 4 * (size_t)n;
  cg.q = pw.vec + 5 * (size_t)n;
  cg.part = pw.part;
  cg.state = state;
  cg.bar = pw.bar;
  cg.err = (int*)(pw.bar + kBarErr);
  cg.tol2 = tol * tol;
  cg.n = n;
  cg.G = pw.persist_G;
  cg.max_iter = max_iter;
  const unsigned lds = (unsigned)(n * sizeof(double));
  if (one_launch && pcg_one_launch(pw, n)) {
    hipLaunchKernelGGL(zero_barrier_kernel, dim3(1), dim3(64), 0, s, pw.bar);
    void* args[] = {&cg};
    RSD_HIP(hipLaunchCooperativeKernel((const void*)cg_persistent_kernel, dim3(cg.G), dim3(kThreads), args, lds, s));
    prof_end(pe, s);
    return RSD_OK;
  }
  if (n <= kPersistMaxN) {
    // one launch per phase over one workgroup per CU; the host checks the
    // status every few iterations and launches past convergence return at once
    if (cg.G == 0) cg.G = pw.cus;
    if (cg.G < 1 || cg.G > pw.part_blk) {
      set_error("pcg: the partial-sum buffer does not hold one slot per workgroup");
      return RSD_ERR_CAPACITY;
    }
    const dim3 grid(cg.G), block(kThreads);
    hipLaunchKernelGGL((cg_phase_kernel<kResidual, double>), grid, block, lds, s, cg, 0);
    hipLaunchKernelGGL((cg_phase_kernel<kPrecond0, float>), grid, block, lds, s, cg, 0);
    int k = 0, chunk = std::max(1, chunk_hint);
    while (true) {
      const int k1 = std::min(max_iter + 1, k + chunk);
      for (; k < k1; ++k) {
        hipLaunchKernelGGL((cg_phase_kernel<kDirection, double>), grid, block, lds, s, cg, k);
        if (k < max_iter) hipLaunchKernelGGL((cg_phase_kernel<kUpdate, float>), grid, block, lds, s, cg, k);
      }
      RSD_HIP(hipGetLastError());
      bool ended = false;
      RSD_TRY(wait_solve_end(pw, slot, s, &ended));
      if (ended || k > max_iter) break;
      chunk = 2;
    }
    prof_end(pe, s);
    return RSD_OK;
  }
  // one launch per phase; the host checks the status every few iterations
  PcgArgs g{};
  g.A = A;
  g.lda = lda;
  g.P = pw.P;
  g.ldp = n;
  g.x = x;
  g.b = b;
  g.r[0] = pw.vec;
  g.r[1] = pw.vec + n;
  g.z = pw.vec + 2 * (size_t)n;
  g.p[0] = pw.vec + 3 * (size_t)n;
  g.p[1] = pw.vec + 4 * (size_t)n;
  g.q = pw.vec + 5 * (size_t)n;
  g.part = pw.part;
  g.state = state;
  g.tol2 = tol * tol;
  g.n = n;
  g.nblk = n / kWaves;
  if (g.nblk > pw.part_blk) {
    set_error("pcg: the partial-sum buffer does not hold one slot per workgroup");
    return RSD_ERR_CAPACITY;
  }
  g.k = 0;
  RSD_TRY(launch_mode<kResidual>(g, s));
  RSD_TRY(launch_mode<kPrecond0>(g, s));
  int k = 0, chunk = std::max(1, chunk_hint);
  while (true) {
    const int k1 = std::min(max_iter + 1, k + chunk);
    for (; k < k1; ++k) {
      g.k = k;
      RSD_TRY(launch_mode<kDirection>(g, s));
      if (k < max_iter) RSD_TRY(launch_mode<kUpdate>(g, s));
    }
    RSD_HIP(hipGetLastError());
    bool ended = false;
    RSD_TRY(wait_solve_end(pw, slot, s, &ended));
    if (ended || k > max_iter) break;
    chunk = 4;
  }
  prof_end(pe, s);
  return RSD_OK;
}

unsigned* pcg_error_word(PcgWork& pw) { return pw.bar ? pw.bar + kBarErr : nullptr; }

}  // namespace rsd
