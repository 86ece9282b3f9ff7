// rsd_free.hip -- free-space grand mobility and the wall image mobility.
//
// Replaces GRPERY_INV_FRI (GRPY_YZD.f90:14-264), the non-periodic grand
// mobility: closed-form self blocks, free-space RPY pair blocks
// (ROTNE_PRAGER_IJ_FTS, stokesian_Green.f90:11-108) or, for overlapping pairs,
// the Rotne-Prager-Yamakawa blocks (correction_method = 1, YAMAKAWA_IJ,
// GRPY_YZD.f90:768-786) or the regularised RPY (correction_method = 2,
// Regular_ROTNE_PRAGER_IJ, Regularization.f90:18-162) in their place; and Wall_GRPERY_INV_FRI
// (Wall_interaction.f90:29-149), the translational mobility of the images in
// the wall z = 0, which STEPPER_Stokesian adds to APP's TT block
// (stepper.f90:67-70).
//
//   free_self_kernel   one thread per particle
//   free_pair_kernel   one lane per pair alpha < beta, 64 pairs to a wave,
//                      the blocks scattered with the reference's sign rules
//   wall_tt_kernel     one thread per 3x3 block (alpha, beta) of TPP
#include <hip/hip_runtime.h>

#include "rsd_free.h"
#include "rsd_scatter.h"

namespace rsd {

using namespace rsdm;

__global__ __launch_bounds__(64) void free_self_kernel(FreeParams p, const double* __restrict__ radii,
                                                       double* __restrict__ G) {
  const int al = blockIdx.x * 64 + threadIdx.x;
  if (al >= p.N) return;
  const double a = radii[al], a3 = a * a * a, sc = p.inv_pimu;
  // ROTNE_PRAGER_TT_SELF / _RR_SELF / _DD_SELF_Y2 (GRPY_YZD.f90:275-315): the
  // Y2 projection of 3/(20 a^3) II is 3/(20 a^3) times the 5x5 identity
  const double tt = sc / (6.0 * a), rr = sc / (8.0 * a3), qq = sc * 3.0 / (20.0 * a3);
  const size_t ld = p.ld;
  const int oT = p.offT + 3 * al, oR = p.offR + 3 * al, oQ = p.offQ + 5 * al;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      store(G, ld, oT + i, oT + j, i == j ? tt : 0.0);
      store(G, ld, oR + i, oR + j, i == j ? rr : 0.0);
      store(G, ld, oT + i, oR + j, 0.0);
      store(G, ld, oR + j, oT + i, 0.0);
    }
  for (int I = 0; I < 5; ++I)
    for (int J = 0; J < 5; ++J) store(G, ld, oQ + I, oQ + J, I == J ? qq : 0.0);
  for (int i = 0; i < 3; ++i)
    for (int J = 0; J < 5; ++J) {
      store(G, ld, oT + i, oQ + J, 0.0);
      store(G, ld, oQ + J, oT + i, 0.0);
      store(G, ld, oR + i, oQ + J, 0.0);
      store(G, ld, oQ + J, oR + i, 0.0);
    }
}

__global__ __launch_bounds__(64) void free_pair_kernel(FreeParams p, const double* __restrict__ pos,
                                                       const double* __restrict__ radii, double* __restrict__ G,
                                                       int* __restrict__ n_overlap) {
  int al, be;
  if (!pair_of_lane((long long)blockIdx.x * 64 + threadIdx.x, p.N, al, be)) return;
  const double ai = radii[al], aj = radii[be];
  // R = p_pos(:,alpha) - p_pos(:,beta), no minimum image (GRPY_YZD.f90:148-152)
  const double x = pos[3 * al] - pos[3 * be], y = pos[3 * al + 1] - pos[3 * be + 1], z = pos[3 * al + 2] - pos[3 * be + 2];
  const double r = sqrt(x * x + y * y + z * z);
  PairOut o;
  pairout_zero(o);
  Radial s;
  if (r <= ai + aj && p.correction_method == 1) {
    // YAMAKAWA_IJ (GRPY_YZD.f90:148-150)
    yamakawa_block(x, y, z, ai, aj, 1.0, o);
    atomicAdd(n_overlap, 1);
  } else if (r <= ai + aj && p.correction_method == 2) {
    // Regular_ROTNE_PRAGER_IJ, kappa = sqrt(pi) / (3 |ai + aj - r|) (Regularization.f90:69)
    const double kap = 1.7724538509055160272981674833411452 / 3.0 / fabs((ai + aj) - r);
    radial_regular(r, kap, s);
    radial_block(x, y, z, s, ai * ai, aj * aj, 1.0, true, 1.0, o);
    atomicAdd(n_overlap, 1);
  } else {
    radial_oseen(r, s);
    radial_block(x, y, z, s, ai * ai, aj * aj, 1.0, true, 1.0, o);
    oseen_quirk(x, y, z, ai * ai, aj * aj, 1.0, o);
  }
  scatter_pair(G, p.ld, p.offT, p.offR, p.offQ, al, be, o, p.inv_pimu);
}

// TT block of the free-space RPY (ROTNE_PRAGER_TT_IJ, stokesian_Green.f90:183-212)
// with its third column negated (Wall_P_Right_tensor, Wall_interaction.f90:176-185).
__device__ __forceinline__ void wall_image_tt(double x, double y, double z, double ai, double aj, double T[3][3]) {
  PairOut o;
  pairout_zero(o);
  Radial s;
  radial_oseen(sqrt(x * x + y * y + z * z), s);
  radial_block(x, y, z, s, ai * ai, aj * aj, 1.0, false, 1.0, o);
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) T[si[q]][sj[q]] = T[sj[q]][si[q]] = o.tt[q];
  for (int i = 0; i < 3; ++i) T[i][2] = -T[i][2];
}

__global__ __launch_bounds__(64) void wall_tt_kernel(int N, double inv_pimu, double zwall, const double* __restrict__ pos,
                                                     const double* __restrict__ radii, double* __restrict__ TPP,
                                                     size_t ld, int accumulate) {
  const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
  if (id >= (long long)N * N) return;
  const int al = (int)(id % N), be = (int)(id / N);  // block (al, be): rows of al, columns of be
  double T[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  bool set = false;
  if (al == be) {
    // self: the particle and its own image, R = (0, 0, 2 z) (Wall_interaction.f90:53-67)
    if (pos[3 * al + 2] <= zwall) {
      wall_image_tt(0.0, 0.0, 2.0 * pos[3 * al + 2], radii[al], radii[al], T);
      set = true;
    }
  } else {
    // the pair (a, b), a < b: particle a and the image of b, R = x_a - (x_b, y_b, -z_b);
    // block (b, a) is the same 3x3 (Wall_interaction.f90:94-122)
    const int a = al < be ? al : be, b = al < be ? be : al;
    if (pos[3 * a + 2] <= zwall || pos[3 * b + 2] <= zwall) {
      wall_image_tt(pos[3 * a] - pos[3 * b], pos[3 * a + 1] - pos[3 * b + 1], pos[3 * a + 2] + pos[3 * b + 2], radii[a],
              radii[b], T);
      set = true;
    }
  }
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      double* e = TPP + (size_t)(3 * al + i) + (size_t)(3 * be + j) * ld;
      const double v = set ? T[i][j] * inv_pimu : 0.0;
      *e = accumulate ? *e + v : v;
    }
}

// MATREV (matrices.f90:4-35) factorises with UPLO = 'U': with the wall's
// non-symmetric TPP added, the grand mobility it inverts is APP's upper
// triangle mirrored.  Copy the TT block's upper triangle onto its lower one.
__global__ __launch_bounds__(64) void mirror_upper_kernel(int n, double* __restrict__ A, size_t ld) {
  const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
  if (id >= (long long)n * n) return;
  const int r = (int)(id % n), c = (int)(id / n);
  if (r > c) A[r + (size_t)c * ld] = A[c + (size_t)r * ld];
}

int free_assemble(const FreeParams& p, const double* d_pos, const double* d_radii, double* G, int* d_overlap,
                  hipStream_t s) {
  hipLaunchKernelGGL(free_self_kernel, dim3((p.N + 63) / 64), dim3(64), 0, s, p, d_radii, G);
  if (p.N > 1) {
    const long long np = (long long)p.N * (p.N - 1) / 2;
    hipLaunchKernelGGL(free_pair_kernel, dim3((unsigned)((np + 63) / 64)), dim3(64), 0, s, p, d_pos, d_radii, G,
                       d_overlap);
  }
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int wall_tt(int N, double mu_f, const double* d_pos, const double* d_radii, double* TPP, size_t ld, bool accumulate,
            hipStream_t s) {
  const long long nb = (long long)N * N;
  // particles within z <= 10 of the wall see its images (Wall_interaction.f90:50,89)
  hipLaunchKernelGGL(wall_tt_kernel, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, N, 1.0 / (kPi * mu_f), 10.0,
                     d_pos, d_radii, TPP, ld, accumulate ? 1 : 0);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int mirror_upper(int n, double* A, size_t ld, hipStream_t s) {
  const long long nb = (long long)n * n;
  hipLaunchKernelGGL(mirror_upper_kernel, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, n, A, ld);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

}  // namespace rsd
