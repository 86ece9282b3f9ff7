// rsd_step.hip -- small per-particle kernels of one Stokesian-dynamics step.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rsd_step.h"

namespace rsd {

// pos_collision_judge, csd_post.f90:383-414: raw (not minimum-image) overlap test.
__global__ void collision_kernel(int N, int Nb, const double* __restrict__ pos, const double* __restrict__ radii,
                                 int* __restrict__ flag) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N - Nb) return;
  const double xa = pos[3 * a], ya = pos[3 * a + 1], za = pos[3 * a + 2], ra = radii[a];
  const int np = N - Nb;
  int hit = 0;
  for (int b = a + 1; b < N; ++b) {
    if (a >= np && b >= np) continue;  // Floc_index 0 on both (boundary particles)
    const double dx = xa - pos[3 * b], dy = ya - pos[3 * b + 1], dz = za - pos[3 * b + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r <= ra + radii[b]) hit = 1;
  }
  if (hit) atomicOr(flag, 1);
}

// PERIOD_CORRECTION, csd_post.f90:65-102 (steady shear)
__device__ __forceinline__ void period_correction(double* R, const StepKernelParams& p) {
  const double lsh = p.LB[2] * p.gamma * p.T;
  if (R[2] > p.LB[2]) {
    R[2] -= p.LB[2];
    R[0] = fmod(R[0] - lsh, p.LB[0]);
  }
  if (R[2] < 0.0) {
    R[2] += p.LB[2];
    R[0] = fmod(R[0] + lsh, p.LB[0]);
  }
  if (R[0] > p.LB[0])
    R[0] -= p.LB[0];
  else if (R[0] < 0.0)
    R[0] += p.LB[0];
  if (R[1] > p.LB[1])
    R[1] -= p.LB[1];
  else if (R[1] < 0.0)
    R[1] += p.LB[1];
}

// STEPPER_conf, stepper.f90:160-227 with calc_uo_bg (csd_post.f90:497-522)
__global__ void euler_kernel(StepKernelParams p, double* __restrict__ pos, const double* __restrict__ upar,
                             double* __restrict__ upos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = p.N;
  if (i >= N) return;
  double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  double u[3];
  for (int r = 0; r < 3; ++r)
    u[r] = upar[3 * i + r] + (p.u_bg[r] + (p.UOE[r] * x[0] + p.UOE[r + 3] * x[1] + p.UOE[r + 6] * x[2]));
  if (i < N - p.Nb) {
    for (int r = 0; r < 3; ++r) x[r] += p.dt * u[r];
    // simplePeriod: the Lees-Edwards wrap and U_pos at the new position
    // (stepper.f90:211-219); otherwise U_pos keeps the pre-step background flow
    if (p.wrap) {
      period_correction(x, p);
      for (int r = 0; r < 3; ++r)
        u[r] = upar[3 * i + r] + (p.u_bg[r] + (p.UOE[r] * x[0] + p.UOE[r + 3] * x[1] + p.UOE[r + 6] * x[2]));
    }
  }
  for (int r = 0; r < 3; ++r) {
    pos[3 * i + r] = x[r];
    upos[3 * i + r] = u[r];
    upos[3 * N + 3 * i + r] = upar[3 * N + 3 * i + r] + p.omega_bg[r];
  }
}

// EIN_transY2 (tensors.f90:905-919): the same 5-vector for every particle
__global__ void fill_ein_kernel(int N, StepEI ei, double* __restrict__ ein) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int q = 0; q < 5; ++q) ein[5 * i + q] = ei.v[q];
}

// source_body_F (source_force.f90:317-342): Fe_z = -gravity on the moving particles
__global__ void fill_fe_kernel(int N, int Nb, double gravity, double* __restrict__ fe) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  fe[3 * i] = 0.0;
  fe[3 * i + 1] = 0.0;
  fe[3 * i + 2] = (i < N - Nb) ? -gravity : 0.0;
  for (int r = 0; r < 3; ++r) fe[3 * N + 3 * i + r] = 0.0;
}

// averageU's force (matrices.f90:230-253): 6 pi mu a_test along x on every particle
__global__ void fill_avgforce_kernel(int N, double f0, double* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * N) return;
  F[i] = (i < 3 * N && i % 3 == 0) ? f0 : 0.0;
}

__global__ void fill_kernel(double* __restrict__ a, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

// yetamu_solve, csd_post.f90:313-367 (SijN_FP = 0 on this path; SijN_B the
// Brownian stresslet, or none)
__global__ __launch_bounds__(256) void yeta_kernel(int N, const double* __restrict__ shyd,
                                                   const double* __restrict__ sb, const double* __restrict__ ein,
                                                   const double* __restrict__ uu, double lv, double mu_f, double phi,
                                                   double* __restrict__ out) {
  __shared__ double s1[256], s2[256], s3[256];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) {
    const double e = -2.0 * ein[4 + 5 * i - 1];
    const double sbi = sb ? sb[4 + 5 * i - 1] : 0.0;
    a += (shyd[4 + 5 * i - 1] + sbi) / e;
    c += sbi / e;
    b += uu[3 * i];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  s3[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
      s3[threadIdx.x] += s3[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0] / lv / mu_f;
    out[1] = phi;
    out[2] = 1.0 / (s2[0] / N);
    out[3] = s3[0] / lv / mu_f;
    out[4] = 0.0;
  }
}

// DMDX_CAL's two displaced configurations (stepper.f90:659-662):
// conf +- (deta/2) z, every particle.
__global__ void brown_conf_kernel(int N, const double* __restrict__ pos, const double* __restrict__ z, double deta,
                                  double* __restrict__ p1, double* __restrict__ p2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * N) return;
  p1[i] = pos[i] + 0.5 * deta * z[i];
  p2[i] = pos[i] - 0.5 * deta * z[i];
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): counter (i, step low, step high, 0) under key
// (seed low, seed high); each 4-word output gives two standard normals by
// the cosine branch of Box-Muller, as GAUSS draws one (stepper.f90:638-648).
__device__ __forceinline__ void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1) {
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0, hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__global__ void brown_normal_kernel(int n, unsigned long long seed, unsigned long long step, double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * i >= n) return;
  unsigned c[4] = {(unsigned)i, (unsigned)step, (unsigned)(step >> 32), 0u};
  philox4x32_10(c, (unsigned)seed, (unsigned)(seed >> 32));
  const double inv32 = 2.3283064365386962890625e-10;  // 2^-32: uniforms in (0, 1)
  for (int q = 0; q < 2 && 2 * i + q < n; ++q) {
    const double a = ((double)c[2 * q] + 0.5) * inv32, b = ((double)c[2 * q + 1] + 0.5) * inv32;
    z[2 * i + q] = sqrt(-2.0 * log(a)) * cos(2.0 * 3.141592653589793238462643383279502884 * b);
  }
}

int launch_collision(int N, int Nb, const double* pos, const double* radii, int* flag, hipStream_t s) {
  if (N - Nb <= 0) return RSD_OK;
  hipLaunchKernelGGL(collision_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, Nb, pos, radii, flag);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_euler(const StepKernelParams& p, double* pos, const double* upar, double* upos, hipStream_t s) {
  hipLaunchKernelGGL(euler_kernel, dim3((p.N + 255) / 256), dim3(256), 0, s, p, pos, upar, upos);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_fill_ein(int N, const StepEI& ei, double* ein, hipStream_t s) {
  hipLaunchKernelGGL(fill_ein_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, ei, ein);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_fill_fe(int N, int Nb, double gravity, double* fe, hipStream_t s) {
  hipLaunchKernelGGL(fill_fe_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, Nb, gravity, fe);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_fill_avgforce(int N, double f0, double* F, hipStream_t s) {
  hipLaunchKernelGGL(fill_avgforce_kernel, dim3((6 * N + 255) / 256), dim3(256), 0, s, N, f0, F);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

__global__ void fill_int_kernel(int* __restrict__ a, size_t n, int v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

int launch_fill_int(int* a, size_t n, int v, hipStream_t s) {
  hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, s, a, n, v);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_fill(double* a, size_t n, double v, hipStream_t s) {
  if (n == 0) return RSD_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, n, v);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_yeta(int N, const double* shyd, const double* sb, const double* ein, const double* uu, double lv,
                double mu_f, double phi, double* out, hipStream_t s) {
  hipLaunchKernelGGL(yeta_kernel, dim3(1), dim3(256), 0, s, N, shyd, sb, ein, uu, lv, mu_f, phi, out);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

// The iterative route's right-hand side [Fe; -EIN] (fe null: zero force)
__global__ void iter_rhs_kernel(int N, const double* __restrict__ fe, const double* __restrict__ ein,
                                double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 11 * N) return;
  b[i] = i < 6 * N ? (fe ? fe[i] : 0.0) : -ein[i - 6 * N];
}

int launch_iter_rhs(int N, const double* fe, const double* ein, double* b, hipStream_t s) {
  hipLaunchKernelGGL(iter_rhs_kernel, dim3((11 * N + 255) / 256), dim3(256), 0, s, N, fe, ein, b);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_brown_conf(int N, const double* pos, const double* z, double deta, double* p1, double* p2, hipStream_t s) {
  hipLaunchKernelGGL(brown_conf_kernel, dim3((3 * N + 255) / 256), dim3(256), 0, s, N, pos, z, deta, p1, p2);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int launch_brown_normal(int n, unsigned long long seed, unsigned long long step, double* z, hipStream_t s) {
  const int pairs = (n + 1) / 2;
  hipLaunchKernelGGL(brown_normal_kernel, dim3((pairs + 255) / 256), dim3(256), 0, s, n, seed, step, z);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

}  // namespace rsd
