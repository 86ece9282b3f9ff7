// rsd_free.h -- launch interface of the free-space and wall mobility kernels.
#pragma once
#include "rsd_internal.h"

namespace rsd {

struct FreeParams {
  double inv_pimu;        // 1/(pi mu_f)
  int N;                  // particles
  int ld;                 // leading dimension of the grand mobility
  int correction_method;  // 0 none, 2 regularised RPY for overlapping pairs
  int offT, offR, offQ;   // rows where particle 0's translation, rotation, stresslet fields start
};

// GRPERY_INV_FRI (GRPY_YZD.f90:14-264) into G (11N x 11N, leading p.ld);
// d_overlap counts the regularised pairs.
int free_assemble(const FreeParams& p, const double* d_pos, const double* d_radii, double* G, int* d_overlap,
                  hipStream_t s);
// Wall_GRPERY_INV_FRI (Wall_interaction.f90:29-149): the 3N x 3N TPP into
// TPP (leading ld), written, or added to it when accumulate (stepper.f90:67-70).
int wall_tt(int N, double mu_f, const double* d_pos, const double* d_radii, double* TPP, size_t ld, bool accumulate,
            hipStream_t s);

// A(r, c) := A(c, r) for r > c (n x n, leading ld): the symmetric matrix
// LAPACK's UPLO = 'U' routines see.
int mirror_upper(int n, double* A, size_t ld, hipStream_t s);

}  // namespace rsd
