// rsd_gnn.hip -- the learned preconditioner of the iterative resistance route:
// the pair bases it weights, the graph network's forward and backward passes,
// its loss and Adam.  The model and the correction it builds are described in
// rsd_gnn.h.
//
//   kblocks_kernel     K's 11x11 blocks on the lubrication structure, one
//                      workgroup per (particle, slot)
//   pair_basis_kernel  the inverse of each pair's 22x22 block by Gauss-Jordan
//                      in LDS (FP64), one workgroup per pair slot
//   gnn_layer_fwd      one workgroup per particle, a lane per hidden unit,
//                      the layer's weights staged in LDS once per workgroup
//   gnn_layer_bwd_*    the same shape backwards; the contributions a particle
//                      receives as a neighbour are gathered through the
//                      reverse slots (no atomics: the gradients are
//                      reproducible)
//   gnn_wgrad          weight gradients as split reductions over the pairs or
//                      the particles, summed in a fixed order
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <vector>

#include "rsd_gnn.h"
#include "rsd_step.h"

namespace rsd {

namespace {

__device__ __forceinline__ int gdof(int N, int g, int c) {
  return c < 3 ? 3 * g + c : (c < 6 ? 3 * N + 3 * g + (c - 3) : 6 * N + 5 * g + (c - 6));
}

// ------------------------------------------------------------ pair bases
__global__ __launch_bounds__(128) void kblocks_kernel(LubBsr<double> L, const double* __restrict__ M, int ld,
                                                      double* __restrict__ kb) {
  const int g = blockIdx.x, slot = blockIdx.y, t = threadIdx.x, N = L.N, nn = L.nnb[g];
  if (slot > nn || t >= 121) return;
  const int d = slot == 0 ? g : L.col[(size_t)g * L.cap + slot - 1];
  const int r = t % 11, c = t / 11, cd = gdof(N, d, c);
  double v;
  if (r < 6) {
    v = (slot == 0 && r == c) ? 1.0 : 0.0;
    const double* b = L.self + 121 * (size_t)g;
    for (int k = 0; k < 11; ++k) v = fma(b[r + 11 * k], M[(size_t)gdof(N, g, k) + (size_t)cd * ld], v);
    for (int q = 0; q < nn; ++q) {
      const int e = L.col[(size_t)g * L.cap + q];
      const double* bk = L.blk + 121 * ((size_t)g * L.cap + q);
      for (int k = 0; k < 11; ++k) v = fma(bk[r + 11 * k], M[(size_t)gdof(N, e, k) + (size_t)cd * ld], v);
    }
  } else {
    v = M[(size_t)gdof(N, g, r) + (size_t)cd * ld];
  }
  kb[((size_t)g * (L.cap + 1) + slot) * 121 + t] = v;
}

__global__ void rev_kernel(LubBsr<double> L, int* __restrict__ rev) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)L.N * L.cap) return;
  const int g = (int)(e / L.cap), q = (int)(e % L.cap);
  if (q >= L.nnb[g]) return;
  const int d = L.col[e];
  int found = -1;
  for (int k = 0; k < L.nnb[d]; ++k)
    if (L.col[(size_t)d * L.cap + k] == g) {
      found = k;
      break;
    }
  rev[e] = found;
}

// inv([K_gg K_gd; K_dg K_dd]) by Gauss-Jordan with partial pivoting; the
// slot's A - D_g and B out
__global__ __launch_bounds__(256) void pair_basis_kernel(LubBsr<double> L, const double* __restrict__ kb,
                                                         const int* __restrict__ rev, const double* __restrict__ dinv,
                                                         double* __restrict__ pa, double* __restrict__ pb) {
  constexpr int n = 22, ne = n * n, per = (ne + 255) / 256;
  __shared__ double A[n][n + 1], X[n][n + 1];
  __shared__ int piv;
  const int e = blockIdx.x, t = threadIdx.x, cap = L.cap;
  const int g = e / cap, q = e % cap;
  if (q >= L.nnb[g]) return;
  const int d = L.col[e], rq = rev[e], cp = cap + 1;
  for (int k = t; k < ne; k += 256) {
    const int r = k % n, c = k / n;
    const int br = r / 11, bc = c / 11, i = r % 11, j = c % 11;
    double v = 0.0;
    if (br == 0 && bc == 0) v = kb[((size_t)g * cp) * 121 + i + 11 * j];
    else if (br == 0) v = kb[((size_t)g * cp + 1 + q) * 121 + i + 11 * j];
    else if (bc == 0) v = rq >= 0 ? kb[((size_t)d * cp + 1 + rq) * 121 + i + 11 * j] : 0.0;
    else v = kb[((size_t)d * cp) * 121 + i + 11 * j];
    A[r][c] = v;
    X[r][c] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      int p = k;
      double best = fabs(A[k][k]);
      for (int i = k + 1; i < n; ++i)
        if (fabs(A[i][k]) > best) {
          best = fabs(A[i][k]);
          p = i;
        }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != k && t < n) {
      double a = A[k][t];
      A[k][t] = A[p][t];
      A[p][t] = a;
      a = X[k][t];
      X[k][t] = X[p][t];
      X[p][t] = a;
    }
    __syncthreads();
    const double inv = 1.0 / A[k][k];
    __syncthreads();
    if (t < n) {
      A[k][t] *= inv;
      X[k][t] *= inv;
    }
    __syncthreads();
    double f[per], fa[per], fx[per];
    for (int u = 0; u < per; ++u) {
      const int kk = t + 256 * u;
      f[u] = fa[u] = fx[u] = 0.0;
      if (kk < ne) {
        f[u] = A[kk % n][k];
        fa[u] = A[k][kk / n];
        fx[u] = X[k][kk / n];
      }
    }
    __syncthreads();
    for (int u = 0; u < per; ++u) {
      const int kk = t + 256 * u;
      if (kk < ne && kk % n != k) {
        A[kk % n][kk / n] -= f[u] * fa[u];
        X[kk % n][kk / n] -= f[u] * fx[u];
      }
    }
    __syncthreads();
  }
  if (t < 121) {
    const int i = t % 11, j = t / 11;
    pa[(size_t)e * 121 + t] = X[i][j] - dinv[(size_t)g * 121 + t];
    pb[(size_t)e * 121 + t] = X[i][11 + j];
  }
}

__global__ void features_kernel(LubBsr<double> L, const double* __restrict__ dist, const double* __restrict__ radii,
                                float* __restrict__ feat, float* __restrict__ xn) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)L.N * L.cap) return;
  const int g = (int)(e / L.cap), q = (int)(e % L.cap);
  if (q == 0) {
    xn[2 * g] = (float)radii[g];
    xn[2 * g + 1] = (float)L.nnb[g] / 16.0f;
  }
  if (q >= L.nnb[g]) return;
  const int d = L.col[e];
  const double ag = radii[g], ad = radii[d], abar = 0.5 * (ag + ad);
  const double h = (dist[e] - ag - ad) / abar;
  feat[4 * e] = (float)h;
  feat[4 * e + 1] = (float)(0.25 * log(fmax(h, 0.0) + 1e-3));
  feat[4 * e + 2] = (float)ag;
  feat[4 * e + 3] = (float)ad;
}

// -------------------------------------------------------------- forward
__global__ void gnn_input_kernel(int N, const float* __restrict__ P, const float* __restrict__ xn,
                                 float* __restrict__ H0) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= N * kGnnH) return;
  const int g = id / kGnnH, i = id % kGnnH;
  const float* W = P + kGnnOffW0;
  const float pre = P[kGnnOffB0 + i] + W[2 * i] * xn[2 * g] + W[2 * i + 1] * xn[2 * g + 1];
  H0[id] = fmaxf(pre, 0.0f);
}

__global__ __launch_bounds__(64) void gnn_layer_fwd_kernel(LubBsr<double> L, const float* __restrict__ P, int l,
                                                           const float* __restrict__ feat,
                                                           const float* __restrict__ Hl, float* __restrict__ Hn,
                                                           float* __restrict__ U, float* __restrict__ AGG,
                                                           float* __restrict__ V) {
  __shared__ float We[kGnnH][kGnnEdgeIn + 1], Wn[kGnnH][kGnnNodeIn + 1];
  __shared__ float in[kGnnEdgeIn], nin[kGnnNodeIn];
  const int g = blockIdx.x, i = threadIdx.x, cap = L.cap, nn = L.nnb[g];
  const float* pwe = P + gnn_off_we(l);
  const float* pwn = P + gnn_off_wn(l);
  for (int k = i; k < kGnnH * kGnnEdgeIn; k += kGnnH) We[k / kGnnEdgeIn][k % kGnnEdgeIn] = pwe[k];
  for (int k = i; k < kGnnH * kGnnNodeIn; k += kGnnH) Wn[k / kGnnNodeIn][k % kGnnNodeIn] = pwn[k];
  const float hg = Hl[(size_t)g * kGnnH + i];
  in[i] = hg;
  const float be = P[gnn_off_be(l) + i];
  float agg = 0.0f;
  __syncthreads();
  for (int q = 0; q < nn; ++q) {
    const size_t e = (size_t)g * cap + q;
    const int d = L.col[e];
    in[kGnnH + i] = Hl[(size_t)d * kGnnH + i];
    if (i < kGnnF) in[2 * kGnnH + i] = feat[4 * e + i];
    __syncthreads();
    float v = be;
    for (int j = 0; j < kGnnEdgeIn; ++j) v = fmaf(We[i][j], in[j], v);
    V[e * kGnnH + i] = v;
    agg += fmaxf(v, 0.0f);
    __syncthreads();
  }
  agg = nn > 0 ? agg / (float)nn : 0.0f;
  AGG[(size_t)g * kGnnH + i] = agg;
  nin[i] = hg;
  nin[kGnnH + i] = agg;
  __syncthreads();
  float u = P[gnn_off_bn(l) + i];
  for (int j = 0; j < kGnnNodeIn; ++j) u = fmaf(Wn[i][j], nin[j], u);
  U[(size_t)g * kGnnH + i] = u;
  Hn[(size_t)g * kGnnH + i] = hg + fmaxf(u, 0.0f);
}

__global__ __launch_bounds__(64) void gnn_out_kernel(LubBsr<double> L, const float* __restrict__ P,
                                                     const float* __restrict__ feat, const float* __restrict__ H3,
                                                     float* __restrict__ OUT) {
  __shared__ float Wo[kGnnEdgeIn], hg[kGnnH];
  const int g = blockIdx.x, t = threadIdx.x, cap = L.cap, nn = L.nnb[g];
  for (int k = t; k < kGnnEdgeIn; k += 64) Wo[k] = P[kGnnOffWo + k];
  hg[t] = H3[(size_t)g * kGnnH + t];
  __syncthreads();
  for (int q = t; q < nn; q += 64) {
    const size_t e = (size_t)g * cap + q;
    const int d = L.col[e];
    float o = P[kGnnOffBo];
    for (int j = 0; j < kGnnH; ++j) o = fmaf(Wo[j], hg[j], o);
    const float* hd = H3 + (size_t)d * kGnnH;
    for (int j = 0; j < kGnnH; ++j) o = fmaf(Wo[kGnnH + j], hd[j], o);
    for (int j = 0; j < kGnnF; ++j) o = fmaf(Wo[2 * kGnnH + j], feat[4 * e + j], o);
    OUT[e] = o;
  }
}

__device__ __forceinline__ double sigmoid(float o) { return 1.0 / (1.0 + exp(-(double)o)); }

__global__ __launch_bounds__(128) void gnn_corr_kernel(LubBsr<double> L, const float* __restrict__ OUT,
                                                       const double* __restrict__ pa, const double* __restrict__ pb,
                                                       double* __restrict__ corr) {
  const int g = blockIdx.x, t = threadIdx.x, cap = L.cap, nn = L.nnb[g];
  if (t >= 121) return;
  double self = 0.0;
  for (int q = 0; q < nn; ++q) {
    const size_t e = (size_t)g * cap + q;
    const double w = sigmoid(OUT[e]);
    self = fma(w, pa[e * 121 + t], self);
    corr[((size_t)g * (cap + 1) + 1 + q) * 121 + t] = w * pb[e * 121 + t];
  }
  corr[(size_t)g * (cap + 1) * 121 + t] = self;
}

// ------------------------------------------------------------- the loss
// out = x with the rows of one field set zeroed: keep_p keeps rows < 6N
__global__ void keep_rows_kernel(int n, int n6, const double* __restrict__ x, double* __restrict__ y, int keep_p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = ((i < n6) == (keep_p != 0)) ? x[i] : 0.0;
}
__global__ void axpby_kernel(int n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                             double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}
// acc[0] += |r|^2, acc[1] += |z|^2 (one workgroup, fixed-order tree)
__global__ __launch_bounds__(256) void sumsq2_kernel(int n, const double* __restrict__ r, const double* __restrict__ z,
                                                     double* __restrict__ acc) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    a = fma(r[i], r[i], a);
    b = fma(z[i], z[i], b);
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[0] += s1[0];
    acc[1] += s2[0];
  }
}
__global__ void zero2_kernel(double* acc) {
  if (threadIdx.x == 0) {
    acc[0] = 0.0;
    acc[1] = 0.0;
  }
}

// dL/do per slot (o the pair weight's logit): scale w (1 - w) sum_z G_g^T
// ((A - D) z_g + B z_d)
__global__ __launch_bounds__(64) void gnn_dout_kernel(LubBsr<double> L, const double* __restrict__ pa,
                                                      const double* __restrict__ pb, const double* __restrict__ G,
                                                      const double* __restrict__ Z, int nz, const double* __restrict__ acc,
                                                      int log_loss, const float* __restrict__ OUT,
                                                      float* __restrict__ dOUT) {
  const int g = blockIdx.x, t = threadIdx.x, cap = L.cap, nn = L.nnb[g], N = L.N;
  const size_t n = 11 * (size_t)N;
  // d(|r|^2 / |z|^2) = 2 r . dr / |z|^2; of the log, divided by |r|^2 / |z|^2 too
  const double scale = log_loss ? 2.0 / acc[0] : 2.0 / acc[1];
  for (int q = t; q < nn; q += 64) {
    const size_t e = (size_t)g * cap + q;
    const int d = L.col[e];
    double sum = 0.0;
    for (int z = 0; z < nz; ++z) {
      const double* Gz = G + z * n;
      const double* Zz = Z + z * n;
      double zg[11], zd[11];
      for (int c = 0; c < 11; ++c) {
        zg[c] = Zz[gdof(N, g, c)];
        zd[c] = Zz[gdof(N, d, c)];
      }
      for (int r = 0; r < 11; ++r) {
        double a = 0.0;
        for (int c = 0; c < 11; ++c) a = fma(pa[e * 121 + r + 11 * c], zg[c], fma(pb[e * 121 + r + 11 * c], zd[c], a));
        sum = fma(Gz[gdof(N, g, r)], a, sum);
      }
    }
    const double w = sigmoid(OUT[e]);
    dOUT[e] = (float)(scale * sum * w * (1.0 - w));
  }
}

// ------------------------------------------------------------- backward
// dH3 from the output layer: g's own slots (first half of Wo) and the slots
// in which g is the neighbour (second half), gathered through rev
__global__ __launch_bounds__(64) void gnn_out_bwd_kernel(LubBsr<double> L, const float* __restrict__ P,
                                                         const int* __restrict__ rev, const float* __restrict__ dOUT,
                                                         float* __restrict__ dH) {
  const int g = blockIdx.x, i = threadIdx.x, cap = L.cap, nn = L.nnb[g];
  const float* Wo = P + kGnnOffWo;
  float dh = 0.0f;
  for (int q = 0; q < nn; ++q) {
    const size_t e = (size_t)g * cap + q;
    dh = fmaf(Wo[i], dOUT[e], dh);
    const int rq = rev[e];
    if (rq >= 0) dh = fmaf(Wo[kGnnH + i], dOUT[(size_t)L.col[e] * cap + rq], dh);
  }
  dH[(size_t)g * kGnnH + i] = dh;
}

// layer l backwards, part one: dU, and dV of g's slots; dHp = dH_{l+1} +
// Wn_h^T dU + We_g^T dV (g's own slots)
__global__ __launch_bounds__(64) void gnn_layer_bwd_a_kernel(LubBsr<double> L, const float* __restrict__ P, int l,
                                                             const float* __restrict__ dHn,
                                                             const float* __restrict__ U,
                                                             const float* __restrict__ V, float* __restrict__ dU,
                                                             float* __restrict__ dV, float* __restrict__ dHp) {
  __shared__ float We[kGnnH][kGnnEdgeIn + 1], Wn[kGnnH][kGnnNodeIn + 1];
  __shared__ float du_s[kGnnH], dv_s[kGnnH];
  const int g = blockIdx.x, i = threadIdx.x, cap = L.cap, nn = L.nnb[g];
  const float* pwe = P + gnn_off_we(l);
  const float* pwn = P + gnn_off_wn(l);
  for (int k = i; k < kGnnH * kGnnEdgeIn; k += kGnnH) We[k / kGnnEdgeIn][k % kGnnEdgeIn] = pwe[k];
  for (int k = i; k < kGnnH * kGnnNodeIn; k += kGnnH) Wn[k / kGnnNodeIn][k % kGnnNodeIn] = pwn[k];
  const float dhn = dHn[(size_t)g * kGnnH + i];
  const float du = U[(size_t)g * kGnnH + i] > 0.0f ? dhn : 0.0f;
  dU[(size_t)g * kGnnH + i] = du;
  du_s[i] = du;
  __syncthreads();
  float dh = dhn, dagg = 0.0f;
  for (int k = 0; k < kGnnH; ++k) {
    dh = fmaf(Wn[k][i], du_s[k], dh);
    dagg = fmaf(Wn[k][kGnnH + i], du_s[k], dagg);
  }
  dagg = nn > 0 ? dagg / (float)nn : 0.0f;
  for (int q = 0; q < nn; ++q) {
    const size_t e = (size_t)g * cap + q;
    const float dv = V[e * kGnnH + i] > 0.0f ? dagg : 0.0f;
    dV[e * kGnnH + i] = dv;
    dv_s[i] = dv;
    __syncthreads();
    for (int k = 0; k < kGnnH; ++k) dh = fmaf(We[k][i], dv_s[k], dh);
    __syncthreads();
  }
  dHp[(size_t)g * kGnnH + i] = dh;
}

// part two: dH_l = dHp + sum over the slots where g is the neighbour of We_d^T dV
__global__ __launch_bounds__(64) void gnn_layer_bwd_b_kernel(LubBsr<double> L, const float* __restrict__ P, int l,
                                                             const int* __restrict__ rev,
                                                             const float* __restrict__ dV,
                                                             const float* __restrict__ dHp, float* __restrict__ dH) {
  __shared__ float We[kGnnH][kGnnH + 1];  // the columns that multiply h_d
  __shared__ float dv_s[kGnnH];
  const int g = blockIdx.x, i = threadIdx.x, cap = L.cap, nn = L.nnb[g];
  const float* pwe = P + gnn_off_we(l);
  for (int k = i; k < kGnnH * kGnnH; k += kGnnH) We[k / kGnnH][k % kGnnH] = pwe[(k / kGnnH) * kGnnEdgeIn + kGnnH + k % kGnnH];
  float dh = dHp[(size_t)g * kGnnH + i];
  __syncthreads();
  for (int q = 0; q < nn; ++q) {
    const size_t e = (size_t)g * cap + q;
    const int rq = rev[e];
    if (rq < 0) continue;  // uniform across the workgroup
    const size_t er = (size_t)L.col[e] * cap + rq;
    dv_s[i] = dV[er * kGnnH + i];
    __syncthreads();
    for (int k = 0; k < kGnnH; ++k) dh = fmaf(We[k][i], dv_s[k], dh);
    __syncthreads();
  }
  dH[(size_t)g * kGnnH + i] = dh;
}

// Edge weight gradients: dW[o][j] = sum_e dY[e][o] in_e[j] (j < 132), the
// bias as j = 132; in_e = [H[g]; H[d]; feat[e]].  Partial sums over chunks of
// particles (grid.y), summed by wgrad_sum_kernel.
__global__ __launch_bounds__(256) void gnn_edge_wgrad_kernel(LubBsr<double> L, int nout, const float* __restrict__ dY,
                                                             const float* __restrict__ H,
                                                             const float* __restrict__ feat, int gchunk,
                                                             float* __restrict__ part) {
  const int id = blockIdx.x * 256 + threadIdx.x, nw = nout * (kGnnEdgeIn + 1);
  if (id >= nw) return;
  const int o = id / (kGnnEdgeIn + 1), j = id % (kGnnEdgeIn + 1), cap = L.cap;
  const int g0 = blockIdx.y * gchunk, g1 = min(L.N, g0 + gchunk);
  float acc = 0.0f;
  for (int g = g0; g < g1; ++g) {
    const int nn = L.nnb[g];
    for (int q = 0; q < nn; ++q) {
      const size_t e = (size_t)g * cap + q;
      float x;
      if (j < kGnnH) x = H[(size_t)g * kGnnH + j];
      else if (j < 2 * kGnnH) x = H[(size_t)L.col[e] * kGnnH + j - kGnnH];
      else if (j < kGnnEdgeIn) x = feat[4 * e + j - 2 * kGnnH];
      else x = 1.0f;
      acc = fmaf(dY[e * nout + o], x, acc);
    }
  }
  part[(size_t)blockIdx.y * nw + id] = acc;
}

// Node weight gradients: dW[o][j] = sum_g dY[g][o] in_g[j], in_g = [A[g]
// (na); B[g] (nb)], bias as j = na + nb.
__global__ __launch_bounds__(256) void gnn_node_wgrad_kernel(int N, int nout, const float* __restrict__ dY,
                                                             const float* __restrict__ A, int na,
                                                             const float* __restrict__ B, int nb, int gchunk,
                                                             float* __restrict__ part) {
  const int ni = na + nb + 1, nw = nout * ni;
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= nw) return;
  const int o = id / ni, j = id % ni;
  const int g0 = blockIdx.y * gchunk, g1 = min(N, g0 + gchunk);
  float acc = 0.0f;
  for (int g = g0; g < g1; ++g) {
    const float x = j < na ? A[(size_t)g * na + j] : (j < na + nb ? B[(size_t)g * nb + j - na] : 1.0f);
    acc = fmaf(dY[(size_t)g * nout + o], x, acc);
  }
  part[(size_t)blockIdx.y * nw + id] = acc;
}

// grad[w_off + o*nin + j] (j < nin) and grad[b_off + o] (j = nin) from the partials
__global__ void wgrad_sum_kernel(int nout, int nin, int chunks, const float* __restrict__ part, float* __restrict__ grad,
                                 int w_off, int b_off) {
  const int nw = nout * (nin + 1);
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nw) return;
  float acc = 0.0f;
  for (int c = 0; c < chunks; ++c) acc += part[(size_t)c * nw + id];
  const int o = id / (nin + 1), j = id % (nin + 1);
  if (j < nin) grad[w_off + o * nin + j] = acc;
  else grad[b_off + o] = acc;
}

__global__ void relu_mask_kernel(int n, const float* __restrict__ dy, const float* __restrict__ y,
                                 float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = y[i] > 0.0f ? dy[i] : 0.0f;
}

__global__ void adam_kernel(int n, float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

constexpr int kChunk = 16;  // particles per partial of a weight-gradient reduction

}  // namespace

// ------------------------------------------------------------------ host
int gnn_model_init(GnnModel& g, unsigned long long seed) {
  std::vector<float> h(kGnnParams, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  auto fill = [&](int off, int rows, int cols, float gain) {
    const float sd = gain * std::sqrt(2.0f / (float)cols);
    for (int k = 0; k < rows * cols; ++k) h[off + k] = sd * nd(rng);
  };
  fill(kGnnOffW0, kGnnH, 2, 1.0f);
  for (int l = 0; l < kGnnLayers; ++l) {
    fill(gnn_off_we(l), kGnnH, kGnnEdgeIn, 1.0f);
    fill(gnn_off_wn(l), kGnnH, kGnnNodeIn, 0.5f);
  }
  // Wo, bo zero: the untrained model weights every pair 1/2
  RSD_HIP(hipGetDevice(&g.dev));
  if (!g.p) RSD_HIP(hipMalloc(&g.p, kGnnParams * sizeof(float)));
  if (!g.m) RSD_HIP(hipMalloc(&g.m, kGnnParams * sizeof(float)));
  if (!g.v) RSD_HIP(hipMalloc(&g.v, kGnnParams * sizeof(float)));
  RSD_HIP(hipMemcpy(g.p, h.data(), kGnnParams * sizeof(float), hipMemcpyHostToDevice));
  RSD_HIP(hipMemset(g.m, 0, kGnnParams * sizeof(float)));
  RSD_HIP(hipMemset(g.v, 0, kGnnParams * sizeof(float)));
  g.t = 0;
  return RSD_OK;
}

void gnn_model_release(GnnModel& g) {
  for (void* p : {(void*)g.p, (void*)g.m, (void*)g.v})
    if (p) (void)hipFree(p);
  g = GnnModel{};
}

int gnn_work_reserve(GnnWork& w, int N, int cap, int nz) {
  if (w.N == N && w.cap == cap && w.nz >= nz) return RSD_OK;
  gnn_work_release(w);
  const size_t S = (size_t)N * cap, n = 11 * (size_t)N;
  const size_t nparts = (size_t)((N + kChunk - 1) / kChunk) * kGnnH * (kGnnEdgeIn + 1);
  RSD_HIP(hipMalloc(&w.kb, (size_t)N * (cap + 1) * 121 * sizeof(double)));
  RSD_HIP(hipMalloc(&w.pa, S * 121 * sizeof(double)));
  RSD_HIP(hipMalloc(&w.pb, S * 121 * sizeof(double)));
  RSD_HIP(hipMalloc(&w.dist, S * sizeof(double)));
  RSD_HIP(hipMalloc(&w.rev, S * sizeof(int)));
  RSD_HIP(hipMalloc(&w.feat, S * kGnnF * sizeof(float)));
  RSD_HIP(hipMalloc(&w.xn, 2 * (size_t)N * sizeof(float)));
  RSD_HIP(hipMalloc(&w.H, 4 * (size_t)N * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.U, 3 * (size_t)N * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.AGG, 3 * (size_t)N * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.V, 3 * S * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.OUT, S * sizeof(float)));
  RSD_HIP(hipMalloc(&w.dOUT, S * sizeof(float)));
  RSD_HIP(hipMalloc(&w.dH, 3 * (size_t)N * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.dU, (size_t)N * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.dV, S * kGnnH * sizeof(float)));
  RSD_HIP(hipMalloc(&w.grad, kGnnParams * sizeof(float)));
  RSD_HIP(hipMalloc(&w.gpart, nparts * sizeof(float)));
  RSD_HIP(hipMalloc(&w.zv, (size_t)nz * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.gv, (size_t)nz * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.wv, 5 * n * sizeof(double)));
  RSD_HIP(hipMalloc(&w.red, 2 * sizeof(double)));
  RSD_HIP(hipHostMalloc(&w.hred, 2 * sizeof(double)));
  w.N = N;
  w.cap = cap;
  w.nz = nz;
  return RSD_OK;
}

void gnn_work_release(GnnWork& w) {
  for (void* p : {(void*)w.kb, (void*)w.pa, (void*)w.pb, (void*)w.dist, (void*)w.rev, (void*)w.feat, (void*)w.xn,
                  (void*)w.H, (void*)w.U, (void*)w.AGG, (void*)w.V, (void*)w.OUT, (void*)w.dOUT, (void*)w.dH,
                  (void*)w.dU, (void*)w.dV, (void*)w.grad, (void*)w.gpart, (void*)w.zv, (void*)w.gv, (void*)w.wv,
                  (void*)w.red})
    if (p) (void)hipFree(p);
  if (w.hred) (void)hipHostFree(w.hred);
  w = GnnWork{};
}

int gnn_prepare(GnnWork& w, const LubBsr<double>& L, const double* M, int ld, const double* radii, const double* dinv,
                hipStream_t s) {
  if (L.N != w.N || L.cap != w.cap) {
    set_error("gnn_prepare: buffers reserved for another size");
    return RSD_ERR_ARG;
  }
  const long long S = (long long)L.N * L.cap;
  hipLaunchKernelGGL(kblocks_kernel, dim3(L.N, L.cap + 1), dim3(128), 0, s, L, M, ld, w.kb);
  hipLaunchKernelGGL(rev_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, L, w.rev);
  hipLaunchKernelGGL(pair_basis_kernel, dim3((unsigned)S), dim3(256), 0, s, L, (const double*)w.kb,
                     (const int*)w.rev, dinv, w.pa, w.pb);
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, L, (const double*)w.dist,
                     radii, w.feat, w.xn);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int gnn_forward(GnnWork& w, const float* params, const LubBsr<double>& L, double* corr, hipStream_t s) {
  const int N = L.N;
  const size_t NH = (size_t)N * kGnnH, SH = (size_t)N * L.cap * kGnnH;
  hipLaunchKernelGGL(gnn_input_kernel, dim3((unsigned)((NH + 255) / 256)), dim3(256), 0, s, N, params,
                     (const float*)w.xn, w.H);
  for (int l = 0; l < kGnnLayers; ++l)
    hipLaunchKernelGGL(gnn_layer_fwd_kernel, dim3(N), dim3(kGnnH), 0, s, L, params, l, (const float*)w.feat,
                       (const float*)(w.H + l * NH), w.H + (l + 1) * NH, w.U + l * NH, w.AGG + l * NH,
                       w.V + l * SH);
  hipLaunchKernelGGL(gnn_out_kernel, dim3(N), dim3(64), 0, s, L, params, (const float*)w.feat,
                     (const float*)(w.H + kGnnLayers * NH), w.OUT);
  hipLaunchKernelGGL(gnn_corr_kernel, dim3(N), dim3(128), 0, s, L, (const float*)w.OUT, (const double*)w.pa,
                     (const double*)w.pb, corr);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

int gnn_loss_grad(GnnWork& w, const float* params, const LubBsr<double>& L, const double* M, int ld,
                  const double* dinv, const double* corr, int nz, unsigned long long seed, unsigned long long draw,
                  bool grad, bool log_loss, double* part, double* loss, hipStream_t s) {
  const int N = L.N, n = 11 * N, n6 = 6 * N;
  if (nz < 1 || nz > w.nz) {
    set_error("gnn_loss_grad: nz outside the reserved buffers");
    return RSD_ERR_ARG;
  }
  const dim3 blk(256), grd((n + 255) / 256);
  double* y = w.wv;
  double* r = w.wv + n;
  double* kw = w.wv + 2 * (size_t)n;  // apply_k's scratch (2n)
  double* t = w.wv + 4 * (size_t)n;
  hipLaunchKernelGGL(zero2_kernel, dim3(1), dim3(64), 0, s, w.red);
  for (int z = 0; z < nz; ++z) {
    double* Zz = w.zv + (size_t)z * n;
    RSD_TRY(launch_brown_normal(n, seed, draw * 4096ull + (unsigned long long)z, Zz, s));
    // r = K P z - z
    RSD_TRY(block_apply<double>(N, dinv, corr, &L, Zz, y, s));
    RSD_TRY(apply_k<double>(L, M, ld, y, r, kw, part, s));
    hipLaunchKernelGGL(axpby_kernel, grd, blk, 0, s, n, 1.0, (const double*)r, -1.0, (const double*)Zz, r);
    hipLaunchKernelGGL(sumsq2_kernel, dim3(1), blk, 0, s, n, (const double*)r, (const double*)Zz, w.red);
    if (!grad) continue;
    // G = K^T r = r_P + M (R_lub r_P + r_Q)
    double* Gz = w.gv + (size_t)z * n;
    hipLaunchKernelGGL(keep_rows_kernel, grd, blk, 0, s, n, n6, (const double*)r, t, 1);
    RSD_TRY(bsr_apply<double>(L, t, y, false, s));
    hipLaunchKernelGGL(keep_rows_kernel, grd, blk, 0, s, n, n6, (const double*)r, kw, 0);
    hipLaunchKernelGGL(axpby_kernel, grd, blk, 0, s, n, 1.0, (const double*)y, 1.0, (const double*)kw, y);
    RSD_TRY(dense_gemv<double>(M, ld, n, y, Gz, part, s));
    hipLaunchKernelGGL(axpby_kernel, grd, blk, 0, s, n, 1.0, (const double*)Gz, 1.0, (const double*)t, Gz);
  }
  RSD_HIP(hipGetLastError());
  if (grad) {
    const size_t NH = (size_t)N * kGnnH, SH = (size_t)N * L.cap * kGnnH;
    const int chunks = (N + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(gnn_dout_kernel, dim3(N), dim3(64), 0, s, L, (const double*)w.pa, (const double*)w.pb,
                       (const double*)w.gv, (const double*)w.zv, nz, (const double*)w.red, log_loss ? 1 : 0,
                       (const float*)w.OUT, w.dOUT);
    // output layer
    hipLaunchKernelGGL(gnn_edge_wgrad_kernel, dim3((kGnnEdgeIn + 1 + 255) / 256, chunks), blk, 0, s, L, 1,
                       (const float*)w.dOUT, (const float*)(w.H + kGnnLayers * NH), (const float*)w.feat, kChunk,
                       w.gpart);
    hipLaunchKernelGGL(wgrad_sum_kernel, dim3((kGnnEdgeIn + 1 + 255) / 256), blk, 0, s, 1, kGnnEdgeIn, chunks,
                       (const float*)w.gpart, w.grad, kGnnOffWo, kGnnOffBo);
    float* dHa = w.dH;           // dH_{l+1}
    float* dHb = w.dH + NH;      // dH_l
    float* dHp = w.dH + 2 * NH;  // partial
    hipLaunchKernelGGL(gnn_out_bwd_kernel, dim3(N), dim3(64), 0, s, L, params, (const int*)w.rev,
                       (const float*)w.dOUT, dHa);
    for (int l = kGnnLayers - 1; l >= 0; --l) {
      hipLaunchKernelGGL(gnn_layer_bwd_a_kernel, dim3(N), dim3(kGnnH), 0, s, L, params, l, (const float*)dHa,
                         (const float*)(w.U + l * NH), (const float*)(w.V + l * SH), w.dU, w.dV, dHp);
      hipLaunchKernelGGL(gnn_layer_bwd_b_kernel, dim3(N), dim3(kGnnH), 0, s, L, params, l, (const int*)w.rev,
                         (const float*)w.dV, (const float*)dHp, dHb);
      const int nwe = kGnnH * (kGnnEdgeIn + 1), nwn = kGnnH * (kGnnNodeIn + 1);
      hipLaunchKernelGGL(gnn_edge_wgrad_kernel, dim3((nwe + 255) / 256, chunks), blk, 0, s, L, kGnnH,
                         (const float*)w.dV, (const float*)(w.H + l * NH), (const float*)w.feat, kChunk, w.gpart);
      hipLaunchKernelGGL(wgrad_sum_kernel, dim3((nwe + 255) / 256), blk, 0, s, kGnnH, kGnnEdgeIn, chunks,
                         (const float*)w.gpart, w.grad, gnn_off_we(l), gnn_off_be(l));
      hipLaunchKernelGGL(gnn_node_wgrad_kernel, dim3((nwn + 255) / 256, chunks), blk, 0, s, N, kGnnH,
                         (const float*)w.dU, (const float*)(w.H + l * NH), kGnnH, (const float*)(w.AGG + l * NH),
                         kGnnH, kChunk, w.gpart);
      hipLaunchKernelGGL(wgrad_sum_kernel, dim3((nwn + 255) / 256), blk, 0, s, kGnnH, kGnnNodeIn, chunks,
                         (const float*)w.gpart, w.grad, gnn_off_wn(l), gnn_off_bn(l));
      std::swap(dHa, dHb);  // dH_l becomes the next (lower) layer's dH_{l+1}
    }
    // input layer: dpre0 = dH0 masked by h0 > 0
    hipLaunchKernelGGL(relu_mask_kernel, dim3((unsigned)((NH + 255) / 256)), blk, 0, s, (int)NH, (const float*)dHa,
                       (const float*)w.H, w.dU);
    const int nw0 = kGnnH * 3;
    hipLaunchKernelGGL(gnn_node_wgrad_kernel, dim3((nw0 + 255) / 256, chunks), blk, 0, s, N, kGnnH,
                       (const float*)w.dU, (const float*)w.xn, 2, (const float*)nullptr, 0, kChunk, w.gpart);
    hipLaunchKernelGGL(wgrad_sum_kernel, dim3((nw0 + 255) / 256), blk, 0, s, kGnnH, 2, chunks, (const float*)w.gpart,
                       w.grad, kGnnOffW0, kGnnOffB0);
    RSD_HIP(hipGetLastError());
  }
  RSD_HIP(hipMemcpyAsync(w.hred, w.red, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipStreamSynchronize(s));
  *loss = w.hred[0] / w.hred[1];
  return RSD_OK;
}

int gnn_adam(GnnModel& g, const float* grad, float lr, hipStream_t s) {
  ++g.t;
  const float c1 = 1.0f - std::pow(0.9f, (float)g.t), c2 = 1.0f - std::pow(0.999f, (float)g.t);
  hipLaunchKernelGGL(adam_kernel, dim3((kGnnParams + 255) / 256), dim3(256), 0, s, kGnnParams, g.p, g.m, g.v, grad,
                     lr, c1, c2);
  RSD_HIP(hipGetLastError());
  return RSD_OK;
}

}  // namespace rsd
