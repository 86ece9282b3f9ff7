// rsd_step.h -- launch interface of the per-particle step kernels.
#pragma once
#include "rsd_internal.h"

namespace rsd {

struct StepKernelParams {
  double LB[3];
  double gamma, T, dt;
  double UOE[9];       // OmegaT + Eij (column-major), csd_post.f90:505
  double u_bg[3], omega_bg[3];
  int N, Nb;
  int wrap;            // simplePeriod: Lees-Edwards wrap after the move
};

struct StepEI {
  double v[5];
};

int launch_collision(int N, int Nb, const double* pos, const double* radii, int* flag, hipStream_t s);
int launch_euler(const StepKernelParams& p, double* pos, const double* upar, double* upos, hipStream_t s);
int launch_fill_ein(int N, const StepEI& ei, double* ein, hipStream_t s);
int launch_fill_fe(int N, int Nb, double gravity, double* fe, hipStream_t s);
int launch_fill_avgforce(int N, double f0, double* F, hipStream_t s);
int launch_fill(double* a, size_t n, double v, hipStream_t s);
int launch_yeta(int N, const double* shyd, const double* ein, const double* uu, double lv, double mu_f, double phi,
                double* out, hipStream_t s);

}  // namespace rsd
