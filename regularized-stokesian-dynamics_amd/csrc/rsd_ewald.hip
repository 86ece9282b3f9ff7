// rsd_ewald.hip -- Ewald-summed RPY grand mobility under Lees-Edwards images.
//
// Replaces PER_GRPERY_INV_FRI (per_GRPY_YZD.f90:13-200) with its calls to
// PER_ROTNE_PRAGER_SELF / PER_ROTNE_PRAGER_IJ (Ewald_summation.f90:19-423)
// and the regularised overlap correction (GRPY_YZD.f90:813-836).
//
// Work decomposition (one wavefront = 64 pairs (alpha, beta..beta+63)):
//   * kweights_kernel   one thread per reciprocal-lattice vector of the
//                       (2NI+1)^3 cube: the 60 wave-vector weights of the
//                       projected k-space block (read back as wave-uniform
//                       scalar loads by the pair kernel).
//   * ewald_self_kernel one wave per particle: self block (cube minus the
//                       origin, k-space, analytic terms).
//   * ewald_pair_kernel one lane per pair, 64 pairs to a wave: skewed minimum image, real-space
//                       image cube, k-space by phase recurrence, optional
//                       overlap correction, then the 18 blocks of the pair
//                       scattered into the 11N x 11N grand mobility with the
//                       reference's sign rules (per_GRPY_YZD.f90:138-188).
#include <hip/hip_runtime.h>

#include "rsd_ewald.h"
#include "rsd_pairmath.h"
#include "rsd_prof.h"
#include "rsd_scatter.h"

namespace rsd {

using namespace rsdm;

// csd_post.f90:24-63 PER_SKEW_CORRECTION
__device__ __forceinline__ double skew_correction(const EwaldParams& p, double& x, double& y, double& z) {
  double X[3];
  for (int i = 0; i < 3; ++i) {
    X[i] = p.LI[i] * x + p.LI[i + 3] * y + p.LI[i + 6] * z;
    X[i] = fmod(X[i], 1.0);
    if (X[i] < 0.0) X[i] += 1.0;
  }
  double RC[3];
  for (int i = 0; i < 3; ++i) RC[i] = p.LR[i] * X[0] + p.LR[i + 3] * X[1] + p.LR[i + 6] * X[2];
  double D = RC[0] * RC[0] + RC[1] * RC[1] + RC[2] * RC[2];
  double bx = RC[0], by = RC[1], bz = RC[2];
  for (int m = 1; m < 8; ++m) {
    const int m1 = (m >> 2) & 1, m2 = (m >> 1) & 1, m3 = m & 1;
    double d0 = RC[0] - (p.LR[0] * m1 + p.LR[3] * m2 + p.LR[6] * m3);
    double d1 = RC[1] - (p.LR[1] * m1 + p.LR[4] * m2 + p.LR[7] * m3);
    double d2 = RC[2] - (p.LR[2] * m1 + p.LR[5] * m2 + p.LR[8] * m3);
    double d = d0 * d0 + d1 * d1 + d2 * d2;
    if (d < D) {
      D = d;
      bx = d0;
      by = d1;
      bz = d2;
    }
  }
  x = bx;
  y = by;
  z = bz;
  return sqrt(D);
}

__global__ void kweights_kernel(EwaldParams p, double* __restrict__ kw) {
  const int side = 2 * p.NI + 1, total = side * side * side;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int nz = idx % side - p.NI, ny = (idx / side) % side - p.NI, nx = idx / (side * side) - p.NI;
  double k[3];
  for (int j = 0; j < 3; ++j)
    k[j] = 2.0 * kPi * (nx * p.LI[0 + 3 * j] + ny * p.LI[1 + 3 * j] + nz * p.LI[2 + 3 * j]);
  kweights(k[0], k[1], k[2], p.lam, 1.0 / p.LV, kw + (size_t)idx * kKW);
  // outside a spherical cutoff the vector weighs nothing (the pair and self
  // kernels then add zeros for it)
  if (p.kmax2 > 0 && nx * nx + ny * ny + nz * nz > p.kmax2)
    for (int i = 0; i < kKW; ++i) kw[(size_t)idx * kKW + i] = 0.0;
}


// one wave per particle: the lanes share the image cube and the wave vectors
// (strided), then sum the self block's tt, rr and qq terms across the wave
__global__ __launch_bounds__(64) void ewald_self_kernel(EwaldParams p, const double* __restrict__ radii,
                                                        const double* __restrict__ kw, double* __restrict__ G) {
  const int al = blockIdx.x, lane = threadIdx.x;
  const double a = radii[al], a2 = a * a;
  PairOut o;
  pairout_zero(o);
  Radial s;
  const int sr = 2 * p.NR + 1, nimg = sr * sr * sr;
  for (int t = lane; t < nimg; t += 64) {  // nz fastest, as nested loops over nx, ny, nz
    const int nx = t / (sr * sr) - p.NR, ny = (t / sr) % sr - p.NR, nz = t % sr - p.NR;
    if (nx == 0 && ny == 0 && nz == 0) continue;
    const double x = p.LR[0] * nx + p.LR[3] * ny + p.LR[6] * nz;
    const double y = p.LR[1] * nx + p.LR[4] * ny + p.LR[7] * nz;
    const double z = p.LR[2] * nx + p.LR[5] * ny + p.LR[8] * nz;
    radial_ewald(sqrt(x * x + y * y + z * z), p.lam, s);
    radial_block(x, y, z, s, a2, a2, p.phix, false, 1.0, o);
  }
  const int side = 2 * p.NI + 1, total = side * side * side;
  for (int q = lane; q < total; q += 64) kspace_accumulate(kw + (size_t)q * kKW, 1.0, 0.0, a2, a2, p.phix, o);
  auto wave_sum = [](double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  };
  for (int q = 0; q < 6; ++q) {
    o.tt[q] = wave_sum(o.tt[q]);
    o.rr[q] = wave_sum(o.rr[q]);
  }
  for (int q = 0; q < 15; ++q) o.qq[q] = wave_sum(o.qq[q]);
  if (lane != 0) return;
  double tt, rr, qq;
  self_analytic(a, p.lam, p.phix, tt, rr, qq);
  const size_t ld = p.ld;
  const int oT = p.offT, oR = p.offR, oQ = p.offQ;
  const double sc = p.inv_pimu;
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    double vt = o.tt[q] + (q < 3 ? tt : 0.0), vr = o.rr[q] + (q < 3 ? rr : 0.0);
    store(G, ld, oT + 3 * al + si[q], oT + 3 * al + sj[q], vt * sc);
    store(G, ld, oT + 3 * al + sj[q], oT + 3 * al + si[q], vt * sc);
    store(G, ld, oR + 3 * al + si[q], oR + 3 * al + sj[q], vr * sc);
    store(G, ld, oR + 3 * al + sj[q], oR + 3 * al + si[q], vr * sc);
  }
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k) {
      double v = (o.qq[k] + (I == J ? qq : 0.0)) * sc;
      store(G, ld, oQ + 5 * al + I, oQ + 5 * al + J, v);
      store(G, ld, oQ + 5 * al + J, oQ + 5 * al + I, v);
    }
  // b, g, h self blocks are zero (per_GRPY_YZD.f90:52-68 fills only a, c, m)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      store(G, ld, oT + 3 * al + i, oR + 3 * al + j, 0.0);
      store(G, ld, oR + 3 * al + j, oT + 3 * al + i, 0.0);
    }
  for (int i = 0; i < 3; ++i)
    for (int J = 0; J < 5; ++J) {
      store(G, ld, oT + 3 * al + i, oQ + 5 * al + J, 0.0);
      store(G, ld, oQ + 5 * al + J, oT + 3 * al + i, 0.0);
      store(G, ld, oR + 3 * al + i, oQ + 5 * al + J, 0.0);
      store(G, ld, oQ + 5 * al + J, oR + 3 * al + i, 0.0);
    }
}

__global__ __launch_bounds__(64) void ewald_pair_kernel(EwaldParams p, const double* __restrict__ pos,
                                                        const double* __restrict__ radii,
                                                        const double* __restrict__ kw, double* __restrict__ G,
                                                        int* __restrict__ n_overlap) {
  // the N(N-1)/2 pairs alpha < beta packed 64 to a wave in row-major order;
  // the last wave's spare lanes repeat the last pair
  int al, be;
  const bool active = pair_of_lane((long long)blockIdx.x * 64 + threadIdx.x, p.N, al, be);
  const int bq = be;
  const double ai = radii[al], aj = radii[bq];
  const double ai2 = ai * ai, aj2 = aj * aj;
  double Rx = pos[3 * al] - pos[3 * bq], Ry = pos[3 * al + 1] - pos[3 * bq + 1], Rz = pos[3 * al + 2] - pos[3 * bq + 2];
  const double dmin = skew_correction(p, Rx, Ry, Rz);

  PairOut o;
  pairout_zero(o);
  Radial s;
  // real space: the full (2NR+1)^3 cube around the minimum image, origin included
  for (int nx = -p.NR; nx <= p.NR; ++nx)
    for (int ny = -p.NR; ny <= p.NR; ++ny)
      for (int nz = -p.NR; nz <= p.NR; ++nz) {
        const double x = Rx + (p.LR[0] * nx + p.LR[3] * ny + p.LR[6] * nz);
        const double y = Ry + (p.LR[1] * nx + p.LR[4] * ny + p.LR[7] * nz);
        const double z = Rz + (p.LR[2] * nx + p.LR[5] * ny + p.LR[8] * nz);
        radial_ewald(sqrt(x * x + y * y + z * z), p.lam, s);
        radial_block(x, y, z, s, ai2, aj2, p.phix, true, 1.0, o);
      }
  // k space: k.R = 2 pi n.(LI R); phases advanced by complex recurrences
  {
    double X[3];
    for (int i = 0; i < 3; ++i) X[i] = p.LI[i] * Rx + p.LI[i + 3] * Ry + p.LI[i + 6] * Rz;
    double wc[3], ws[3], pc[3], ps[3];
    for (int d = 0; d < 3; ++d) {
      sincos(2.0 * kPi * X[d], &ws[d], &wc[d]);
      sincos(-2.0 * kPi * X[d] * p.NI, &ps[d], &pc[d]);  // w^{-NI}
    }
    // Every k-space term is even in k (cosine terms with even weights, sine
    // terms with odd ones; k = 0 weighs nothing), so the cube's sum is twice
    // the sum over its positive half: n1 > 0, or n1 = 0 and n2 > 0, or
    // n1 = n2 = 0 and n3 > 0.  The phases start at 2 for the factor.
    const int NI = p.NI, side = 2 * NI + 1;
    double z1c = 2.0, z1s = 0.0;
    for (int n1 = 0; n1 <= NI; ++n1) {
      double z2c = z1c, z2s = z1s;  // n2 from 0 when n1 = 0, else from -NI
      if (n1 > 0) {
        z2c = z1c * pc[1] - z1s * ps[1];
        z2s = z1c * ps[1] + z1s * pc[1];
      }
      for (int n2 = (n1 == 0 ? 0 : -NI); n2 <= NI; ++n2) {
        int n3 = -NI;
        double zc = z2c * pc[2] - z2s * ps[2], zs = z2c * ps[2] + z2s * pc[2];
        if (n1 == 0 && n2 == 0) {  // n3 from 1
          n3 = 1;
          zc = z2c * wc[2] - z2s * ws[2];
          zs = z2c * ws[2] + z2s * wc[2];
        }
        const double* kwrow = kw + (size_t)(((n1 + NI) * side + (n2 + NI)) * side + NI) * kKW;  // n3 = 0
        for (; n3 <= NI; ++n3) {
          kspace_accumulate(kwrow + (ptrdiff_t)n3 * kKW, zc, zs, ai2, aj2, p.phix, o);
          const double tc = zc * wc[2] - zs * ws[2];
          zs = zc * ws[2] + zs * wc[2];
          zc = tc;
        }
        const double tc = z2c * wc[1] - z2s * ws[1];
        z2s = z2c * ws[1] + z2s * wc[1];
        z2c = tc;
      }
      const double tc = z1c * wc[0] - z1s * ws[0];
      z1s = z1c * ws[0] + z1s * wc[0];
      z1c = tc;
    }
  }
  if (!active) return;
  if (p.correction_method != 0 && dmin <= ai + aj) {
    // per_GRPY_YZD.f90:98-116: 1 Yamakawa, 2 regularised
    if (p.correction_method == 1)
      overlap_yamakawa_correction(Rx, Ry, Rz, ai, aj, o);
    else
      overlap_regular_correction(Rx, Ry, Rz, ai, aj, o);
    atomicAdd(n_overlap, 1);
  }
  // scatter (per_GRPY_YZD.f90:122-188), scaled by 1/(pi mu_f)
  scatter_pair(G, p.ld, p.offT, p.offR, p.offQ, al, be, o, p.inv_pimu);
}

int ewald_assemble(const EwaldParams& p, const double* d_pos, const double* d_radii, double* d_kw, double* G,
                   int* d_overlap, hipStream_t s, bool profiled) {
  const int side = 2 * p.NI + 1, total = side * side * side;
  // algorithmic work convention: 1000 flop per real-space and 250 per k-space pair-image
  const double nr3 = (double)(2 * p.NR + 1) * (2 * p.NR + 1) * (2 * p.NR + 1), nk = (double)total - 1.0;
  hipEvent_t pe = profiled ? prof_begin(kProfEwaldSelf, p.N * (1000.0 * (nr3 - 1.0) + 250.0 * nk), 0.0, s) : nullptr;
  hipLaunchKernelGGL(kweights_kernel, dim3((total + 63) / 64), dim3(64), 0, s, p, d_kw);
  hipLaunchKernelGGL(ewald_self_kernel, dim3(p.N), dim3(64), 0, s, p, d_radii, (const double*)d_kw, G);
  prof_end(pe, s);
  if (p.N > 1) {
    const long long np = (long long)p.N * (p.N - 1) / 2;
    dim3 grid((unsigned)((np + 63) / 64));
    const double pairs = 0.5 * (double)p.N * (p.N - 1);
    hipEvent_t pp =
        profiled ? prof_begin(kProfEwaldPair, pairs * (1000.0 * nr3 + 250.0 * nk), pairs * 242.0 * 8.0, s) : nullptr;
    hipLaunchKernelGGL(ewald_pair_kernel, grid, dim3(64), 0, s, p, d_pos, d_radii, (const double*)d_kw, G, d_overlap);
    prof_end(pp, s);
  }
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

}  // namespace rsd
