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
int launch_fill_int(int* a, size_t n, int v, hipStream_t s);
// yetamu_solve; sb (the Brownian stresslet SijN_B) may be null
int launch_yeta(int N, const double* shyd, const double* sb, const double* ein, const double* uu, double lv,
                double mu_f, double phi, double* out, hipStream_t s);
// [Fe; -EIN] (11N) for the iterative route; fe null: zero force
int launch_iter_rhs(int N, const double* fe, const double* ein, double* b, hipStream_t s);
// DMDX_CAL's displaced configurations p1, p2 = pos +- (deta/2) z
int launch_brown_conf(int N, const double* pos, const double* z, double deta, double* p1, double* p2, hipStream_t s);
// n standard normals from the counter-based generator (seed, step)
int launch_brown_normal(int n, unsigned long long seed, unsigned long long step, double* z, hipStream_t s);

}  // namespace rsd
