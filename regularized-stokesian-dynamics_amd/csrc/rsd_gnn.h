// rsd_gnn.h -- the learned preconditioner of the iterative resistance route
// (rsd_gnn.hip).  No counterpart in the reference, whose GMRES
// (gmres_mod_gmres.f90:23-61) is unpreconditioned.
//
// The preconditioner is block-Jacobi (the inverses D_g of K's 11x11 diagonal
// blocks) plus, for each neighbour pair (g, d) of the lubrication structure,
// weighted pieces of the inverse of K's 22x22 pair block
//     inv([K_gg K_gd; K_dg K_dd]) = [A_gd B_gd; . .]:
//     P_gg = D_g + sum_d w(g,d) (A_gd - D_g),   P_gd = w(g,d) B_gd,
// a weighted additive-Schwarz correction over pairs.  A graph network predicts
// each pair's weight w in (0, 1) (a sigmoid) from rotation-invariant features:
// the gap, the radii and, through three rounds of message passing, the
// neighbourhood (how many close pairs share the particle).  w -> 0 is
// block-Jacobi; for an isolated pair w = 1 is the exact inverse; the untrained
// model (zero output weights) gives every pair w = 1/2.
//
// The network (FP32, hidden width 64):
//   x_g = [a_g, n_g / 16]                         (radius, neighbour count)
//   f_gd = [h/abar, log(max(h/abar, 0) + 1e-3)/4, a_g, a_d]  (h the gap)
//   h0_g = relu(W0 x_g + b0)
//   three layers l:  m_gd = relu(We_l [h_g; h_d; f_gd] + be_l),
//                    s_g = mean_d m_gd,
//                    h_g <- h_g + relu(Wn_l [h_g; s_g] + bn_l)
//   w(g,d) = sigmoid(Wo [h_g; h_d; f_gd] + bo)
// It is trained on the systems the step itself assembles, to minimise the
// stochastic Frobenius loss sum_z |K P z - z|^2 / sum_z |z|^2 over Gaussian
// z, with hand-written backward kernels and Adam.
#pragma once
#include "rsd_iter.h"

namespace rsd {

constexpr int kGnnH = 64;                          // hidden width
constexpr int kGnnF = 4;                           // pair features
constexpr int kGnnLayers = 3;
constexpr int kGnnEdgeIn = 2 * kGnnH + kGnnF;      // 132
constexpr int kGnnNodeIn = 2 * kGnnH;              // 128
constexpr int kGnnOffW0 = 0;                       // W0 [64][2] row-major
constexpr int kGnnOffB0 = 2 * kGnnH;
constexpr int kGnnLayerBase = 3 * kGnnH;
constexpr int kGnnLayerSize = kGnnH * kGnnEdgeIn + kGnnH + kGnnH * kGnnNodeIn + kGnnH;  // We be Wn bn
constexpr int kGnnOffWo = kGnnLayerBase + kGnnLayers * kGnnLayerSize;  // Wo [132]
constexpr int kGnnOffBo = kGnnOffWo + kGnnEdgeIn;
constexpr int kGnnParams = kGnnOffBo + 1;
__host__ __device__ constexpr int gnn_off_we(int l) { return kGnnLayerBase + l * kGnnLayerSize; }
__host__ __device__ constexpr int gnn_off_be(int l) { return gnn_off_we(l) + kGnnH * kGnnEdgeIn; }
__host__ __device__ constexpr int gnn_off_wn(int l) { return gnn_off_be(l) + kGnnH; }
__host__ __device__ constexpr int gnn_off_bn(int l) { return gnn_off_wn(l) + kGnnH * kGnnNodeIn; }

// A model: parameters and Adam moments on the device.
struct GnnModel {
  int dev = -1;
  float* p = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  long long t = 0;  // Adam steps taken
};
int gnn_model_init(GnnModel& g, unsigned long long seed);
void gnn_model_release(GnnModel& g);

// Per-suspension buffers (N particles, cap neighbour slots each).
struct GnnWork {
  int N = 0, cap = 0, nz = 0;
  double* kb = nullptr;    // K's blocks on the structure: slot 0 self, 1+q neighbour q (121 each)
  double* pa = nullptr;    // A_gd - D_g per slot
  double* pb = nullptr;    // B_gd per slot
  double* dist = nullptr;  // pair distance per slot (written by the lubrication)
  int* rev = nullptr;      // slot of g in d's list, per slot (g, q)
  float* feat = nullptr;   // f_gd per slot
  float* xn = nullptr;     // x_g
  float* H = nullptr;      // h after the input layer and each layer: 4 x N x 64
  float* U = nullptr;      // node pre-activations: 3 x N x 64
  float* AGG = nullptr;    // s_g: 3 x N x 64
  float* V = nullptr;      // edge pre-activations: 3 x N cap x 64
  float* OUT = nullptr;    // the pair weight's logit per slot
  float* dOUT = nullptr;
  float* dH = nullptr;     // 3 x N x 64 (two ping-pong + a partial)
  float* dU = nullptr;     // N x 64
  float* dV = nullptr;     // N cap x 64
  float* grad = nullptr;   // kGnnParams
  float* gpart = nullptr;  // split reductions
  double* zv = nullptr;    // nz Gaussian vectors (11N each)
  double* gv = nullptr;    // nz vectors K^T (K P z - z)
  double* wv = nullptr;    // work: 5 vectors
  double* red = nullptr;   // reductions (2)
  double* hred = nullptr;  // pinned copy
};
int gnn_work_reserve(GnnWork& w, int N, int cap, int nz);
void gnn_work_release(GnnWork& w);

// K's blocks on the lubrication structure, the reverse slots, the pair
// inverses' pieces and the features (dinv: block-Jacobi's inverses).
int gnn_prepare(GnnWork& w, const LubBsr<double>& L, const double* M, int ld, const double* radii, const double* dinv,
                hipStream_t s);
// The network's pair weights and the correction blocks (corr: the layout
// block_apply reads, slot 0 self and 1+q neighbour q, 121 each).
int gnn_forward(GnnWork& w, const float* params, const LubBsr<double>& L, double* corr, hipStream_t s);
// The loss for nz Gaussian vectors drawn from (seed, draw) and, when grad,
// its gradient with respect to the parameters into w.grad (log_loss: of
// log(loss)).  Needs gnn_forward's activations; part: dense_gemv's scratch.
int gnn_loss_grad(GnnWork& w, const float* params, const LubBsr<double>& L, const double* M, int ld,
                  const double* dinv, const double* corr, int nz, unsigned long long seed, unsigned long long draw,
                  bool grad, bool log_loss, double* part, double* loss, hipStream_t s);
// One Adam step of the model with w.grad.
int gnn_adam(GnnModel& g, const float* grad, float lr, hipStream_t s);

}  // namespace rsd
