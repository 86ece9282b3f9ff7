// rsd_lub.hip -- pairwise lubrication resistance (lubmxcalc, hydro_lub.f90:22-308).
//
// One wavefront per particle gamma; its lanes sweep every partner delta.
// Each pair is oriented as the reference orients it (lower index minus
// higher index, then PER_SKEW_CORRECTION), so the (1,1)/(2,2) self blocks
// and the (1,2)/(2,1) cross blocks get the reference's signs.  Self blocks
// of gamma are reduced across the wave in registers and written by their
// owning wave; cross blocks of (alpha<beta) are written by alpha's wave.
// No atomics on the matrix, so the result is bitwise repeatable.
#include <hip/hip_runtime.h>

#include "rsd_lub.h"
#include "rsd_lubmath.h"
#include "rsd_prof.h"

namespace rsd {

using namespace rsdm;

__device__ __forceinline__ double lub_skew(const LubParams& p, double& x, double& y, double& z) {
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

__device__ __forceinline__ void addm(double* A, int ld, int r, int c, double v) {
  A[(size_t)r + (size_t)c * ld] += v;
}

constexpr int kSelfAcc = 6 + 6 + 3 + 15 + 15 + 15;  // a, c, b, g, h, m

__global__ __launch_bounds__(64) void lub_kernel(LubParams p, const double* __restrict__ pos,
                                                 const double* __restrict__ radii, const LubTables* __restrict__ tab,
                                                 int* __restrict__ contacts) {
  const int gm = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = p.N;
  const double sc = p.scale;
  double acc[kSelfAcc];
  for (int q = 0; q < kSelfAcc; ++q) acc[q] = 0.0;
  for (int dl = lane; dl < N; dl += 64) {
    if (dl == gm) continue;
    const int lo = gm < dl ? gm : dl, hi = gm < dl ? dl : gm;
    if (lo > N - p.Nb - 1) continue;  // hydro_lub.f90:68-71
    double rx = pos[3 * lo] - pos[3 * hi], ry = pos[3 * lo + 1] - pos[3 * hi + 1], rz = pos[3 * lo + 2] - pos[3 * hi + 2];
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (p.is_period) r = lub_skew(p, rx, ry, rz);
    const double r_ave = 0.5 * (radii[lo] + radii[hi]);
    double s = r / r_ave;
    if (s <= 2.0) {
      s = 4.00001 - s;  // contact reflection, hydro_lub.f90:85-92
      if (gm == lo) atomicAdd(contacts, 1);
    } else if (s > 4.0) {
      continue;
    }
    double v[19];
    lubcalc(s, *tab, v);
    const double e[3] = {rx / r, ry / r, rz / r};
    YProj yp;
    yproj(e[0], e[1], e[2], yp);
    // self contribution of gamma: (1,1) block if gamma is the lower index, else (2,2)
    const double sg = (gm == lo) ? 1.0 : -1.0;
    LubBlock b11;
    lub_block(e, v[0], v[2], v[4], v[6], v[8], v[10], v[12], v[14], yp, b11);
    double mm[15];
    lub_mmx(v[16], v[17], v[18], yp, mm);
    for (int q = 0; q < 6; ++q) {
      acc[q] += b11.a[q];
      acc[6 + q] += b11.c[q];
    }
    for (int k = 0; k < 3; ++k) acc[12 + k] += sg * b11.b[k];
    for (int q = 0; q < 15; ++q) {
      acc[15 + q] += sg * b11.g[q];
      acc[30 + q] += b11.h[q];
      acc[45 + q] += mm[q];
    }
    if (gm != lo) continue;
    // cross blocks of (alpha=gm, beta=dl), written once by the lower index
    LubBlock b12;
    lub_block(e, v[1], v[3], v[5], v[7], v[9], v[11], v[13], v[15], yp, b12);
    const int al = gm, be = dl;
    const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) {
      const int i = si[q], j = sj[q];
      const double av = sc * b12.a[q], cv = sc * b12.c[q];
      addm(p.PP, p.ldPP, 3 * al + i, 3 * be + j, av);
      addm(p.PP, p.ldPP, 3 * be + i, 3 * al + j, av);
      addm(p.PP, p.ldPP, 3 * N + 3 * al + i, 3 * N + 3 * be + j, cv);
      addm(p.PP, p.ldPP, 3 * N + 3 * be + i, 3 * N + 3 * al + j, cv);
      if (i != j) {
        addm(p.PP, p.ldPP, 3 * al + j, 3 * be + i, av);
        addm(p.PP, p.ldPP, 3 * be + j, 3 * al + i, av);
        addm(p.PP, p.ldPP, 3 * N + 3 * al + j, 3 * N + 3 * be + i, cv);
        addm(p.PP, p.ldPP, 3 * N + 3 * be + j, 3 * N + 3 * al + i, cv);
      }
    }
    // B12(i,j) = eps_ijk b_k : RT(alpha,beta)=B12, RT(beta,alpha)=-B12, TR = RT^T
    const double b0 = sc * b12.b[0], b1 = sc * b12.b[1], b2 = sc * b12.b[2];
    const double B[3][3] = {{0.0, b2, -b1}, {-b2, 0.0, b0}, {b1, -b0, 0.0}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        if (i == j) continue;
        addm(p.PP, p.ldPP, 3 * N + 3 * al + i, 3 * be + j, B[i][j]);
        addm(p.PP, p.ldPP, 3 * be + j, 3 * N + 3 * al + i, B[i][j]);
        addm(p.PP, p.ldPP, 3 * N + 3 * be + i, 3 * al + j, -B[i][j]);
        addm(p.PP, p.ldPP, 3 * al + j, 3 * N + 3 * be + i, -B[i][j]);
      }
    for (int I = 0; I < 5; ++I)
      for (int c = 0; c < 3; ++c) {
        const double g = sc * b12.g[I + 5 * c], h = sc * b12.h[I + 5 * c];
        addm(p.PQ, p.ldPQ, 3 * be + c, 5 * al + I, g);
        addm(p.PQ, p.ldPQ, 3 * al + c, 5 * be + I, -g);
        addm(p.PQ, p.ldPQ, 3 * N + 3 * be + c, 5 * al + I, h);
        addm(p.PQ, p.ldPQ, 3 * N + 3 * al + c, 5 * be + I, h);
      }
    int k = 0;
    for (int I = 0; I < 5; ++I)
      for (int J = I; J < 5; ++J, ++k) {
        const double m = sc * mm[k];
        addm(p.QQ, p.ldQQ, 5 * al + I, 5 * be + J, m);
        addm(p.QQ, p.ldQQ, 5 * be + I, 5 * al + J, m);
        if (J != I) {
          addm(p.QQ, p.ldQQ, 5 * al + J, 5 * be + I, m);
          addm(p.QQ, p.ldQQ, 5 * be + J, 5 * al + I, m);
        }
      }
  }
  // fixed-order wave reduction of the self accumulators
  for (int q = 0; q < kSelfAcc; ++q) {
    double x = acc[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    acc[q] = x;
  }
  if (lane != 0) return;
  const int g = gm;
  const int si[6] = {0, 1, 2, 0, 0, 1}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    const int i = si[q], j = sj[q];
    addm(p.PP, p.ldPP, 3 * g + i, 3 * g + j, sc * acc[q]);
    addm(p.PP, p.ldPP, 3 * N + 3 * g + i, 3 * N + 3 * g + j, sc * acc[6 + q]);
    if (i != j) {
      addm(p.PP, p.ldPP, 3 * g + j, 3 * g + i, sc * acc[q]);
      addm(p.PP, p.ldPP, 3 * N + 3 * g + j, 3 * N + 3 * g + i, sc * acc[6 + q]);
    }
  }
  const double b0 = sc * acc[12], b1 = sc * acc[13], b2 = sc * acc[14];
  const double B[3][3] = {{0.0, b2, -b1}, {-b2, 0.0, b0}, {b1, -b0, 0.0}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (i == j) continue;
      addm(p.PP, p.ldPP, 3 * N + 3 * g + i, 3 * g + j, B[i][j]);
      addm(p.PP, p.ldPP, 3 * g + j, 3 * N + 3 * g + i, B[i][j]);
    }
  for (int I = 0; I < 5; ++I)
    for (int c = 0; c < 3; ++c) {
      addm(p.PQ, p.ldPQ, 3 * g + c, 5 * g + I, sc * acc[15 + I + 5 * c]);
      addm(p.PQ, p.ldPQ, 3 * N + 3 * g + c, 5 * g + I, sc * acc[30 + I + 5 * c]);
    }
  int k = 0;
  for (int I = 0; I < 5; ++I)
    for (int J = I; J < 5; ++J, ++k) {
      addm(p.QQ, p.ldQQ, 5 * g + I, 5 * g + J, sc * acc[45 + k]);
      if (J != I) addm(p.QQ, p.ldQQ, 5 * g + J, 5 * g + I, sc * acc[45 + k]);
    }
}

int lub_add(const LubParams& p, const double* d_pos, const double* d_radii, const LubTables* d_tab, int* d_contacts,
            hipStream_t s) {
  if (p.N < 1) return RSD_OK;
  hipEvent_t pe = prof_begin(kProfLub, 0.0, 0.0, s);
  hipLaunchKernelGGL(lub_kernel, dim3(p.N), dim3(64), 0, s, p, d_pos, d_radii, d_tab, d_contacts);
  prof_end(pe, s);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

}  // namespace rsd
