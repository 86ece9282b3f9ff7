// rsd_prof.h -- optional per-kernel-family timing with HIP events recorded on
// the launching stream (enabled by rsd_profile_enable; off by default).
#pragma once
#include <hip/hip_runtime.h>

namespace rsd {

enum ProfFamily : int {
  kProfGemm = 0,      // gemm64_kernel (all instantiations)
  kProfPotrfDiag = 1, // potrf_dag_kernel (one dataflow Cholesky per launch)
  kProfTrsv = 2,      // potrs_fwd_kernel + potrs_bwd_kernel (triangular solves)
  kProfEwaldPair = 3, // ewald_pair_kernel
  kProfEwaldSelf = 4, // ewald_self_kernel + kweights_kernel
  kProfLub = 5,       // lub_kernel
  kProfGemv = 6,      // gemv kernels
  kProfPcg = 7,       // pcg_kernel (one conjugate-gradient solve per record)
  kProfGemvSplit64 = 8,  // gemv_split_kernel<double> + its sum (the iterative route's M products, FP64)
  kProfGemvSplit32 = 9,  // gemv_split_kernel<float> + its sum (the mixed preconditioner's, FP32)
  kProfInnerGraph = 10,  // one replay of the captured FP32 inner GMRES (the mixed preconditioner)
  kProfFamilies = 11
};

bool prof_on();
// returns an event to record before the launch (or nullptr when profiling is off)
hipEvent_t prof_begin(int family, double flops, double bytes, hipStream_t s);
void prof_end(hipEvent_t begin, hipStream_t s);

}  // namespace rsd
