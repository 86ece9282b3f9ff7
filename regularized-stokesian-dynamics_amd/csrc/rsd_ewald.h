// rsd_ewald.h -- launch interface of the Ewald grand-mobility kernels.
#pragma once
#include "rsd_internal.h"

namespace rsd {

// Kernel-argument copy of the lattice state the reference keeps in the
// LATTICE_BASE module (mymodules.f90:33-41) plus the assembly constants.
struct EwaldParams {
  double LR[9], LI[9];  // column-major 3x3 lattice and its inverse
  double LV, lam, phix;  // cell volume, splitting parameter, 1 - phi/5
  double inv_pimu;       // 1/(pi mu_f)
  int NR, NI;            // real / reciprocal image half-widths
  int N;                 // particles
  int ld;                // leading dimension of the grand mobility
  int correction_method; // 0 none, 2 regularised overlap correction
  // rows (and columns) where the translation, rotation and stresslet fields
  // of particle 0 start: 0, 3N, 6N in the reference's order
  // (per_GRPY_YZD.f90:165-184); other orders place the same blocks elsewhere
  int offT, offR, offQ;
  // > 0: only reciprocal vectors n with |n|^2 <= kmax2 (a spherical cutoff in
  // the (2NI+1)^3 cube); 0: the whole cube, as the reference sums it
  int kmax2;
};

// Assemble the 11N x 11N grand mobility into G (leading dimension p.ld).
// d_kw must hold (2NI+1)^3 * 64 doubles; d_overlap counts corrected pairs.
// With profiled = false the launches bypass the kernel-family timers (which
// create HIP events and grow a host list when timing is on).
int ewald_assemble(const EwaldParams& p, const double* d_pos, const double* d_radii, double* d_kw, double* G,
                   int* d_overlap, hipStream_t s, bool profiled = true);

}  // namespace rsd
