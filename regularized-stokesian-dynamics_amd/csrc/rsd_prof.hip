// rsd_prof.hip -- per-kernel-family HIP-event timing (see rsd_prof.h).
#include "rsd_prof.h"

#include <vector>

#include "rsd_internal.h"

namespace rsd {

namespace {
struct Rec {
  hipEvent_t a, b;
  int fam;
  double flops, bytes;
};
struct Prof {
  unsigned mask = 0;  // families timed
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t next = 0;
  long launches[kProfFamilies] = {};
  double ms[kProfFamilies] = {}, flops[kProfFamilies] = {}, bytes[kProfFamilies] = {};
};
Prof& P() {
  static Prof p;
  return p;
}
hipEvent_t take() {
  Prof& p = P();
  if (p.next == p.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p.pool.push_back(e);
  }
  return p.pool[p.next++];
}
}  // namespace

bool prof_on() { return P().mask != 0; }

hipEvent_t prof_begin(int family, double flops, double bytes, hipStream_t s) {
  Prof& p = P();
  if (!((p.mask >> family) & 1u) || p.recs.size() >= 200000) return nullptr;
  // launches being captured into a HIP graph are timed as the graph's launch
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return nullptr;
  hipEvent_t a = take(), b = take();
  if (!a || !b) return nullptr;
  (void)hipEventRecord(a, s);
  p.recs.push_back({a, b, family, flops, bytes});
  return a;
}

void prof_end(hipEvent_t begin, hipStream_t s) {
  if (!begin) return;
  Prof& p = P();
  for (size_t i = p.recs.size(); i-- > 0;)
    if (p.recs[i].a == begin) {
      (void)hipEventRecord(p.recs[i].b, s);
      return;
    }
}

static int drain() {
  Prof& p = P();
  for (const Rec& r : p.recs) {
    RSD_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    RSD_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    p.launches[r.fam]++;
    p.ms[r.fam] += ms;
    p.flops[r.fam] += r.flops;
    p.bytes[r.fam] += r.bytes;
  }
  p.recs.clear();
  p.next = 0;
  return RSD_OK;
}

}  // namespace rsd

extern "C" {

int rsd_profile_enable_families(unsigned mask) {
  rsd::Prof& p = rsd::P();
  int rc = rsd::drain();
  for (int f = 0; f < rsd::kProfFamilies; ++f) {
    p.launches[f] = 0;
    p.ms[f] = p.flops[f] = p.bytes[f] = 0.0;
  }
  p.mask = mask & ((1u << rsd::kProfFamilies) - 1u);
  return rc;
}

int rsd_profile_enable(int on) { return rsd_profile_enable_families(on ? ~0u : 0u); }

int rsd_profile_read(int family, long* launches, double* ms, double* flops, double* bytes) {
  if (family < 0 || family >= rsd::kProfFamilies) return RSD_ERR_ARG;
  int rc = rsd::drain();
  if (rc) return rc;
  rsd::Prof& p = rsd::P();
  if (launches) *launches = p.launches[family];
  if (ms) *ms = p.ms[family];
  if (flops) *flops = p.flops[family];
  if (bytes) *bytes = p.bytes[family];
  return RSD_OK;
}

const char* rsd_profile_family_name(int family) {
  static const char* names[] = {"gemm64_kernel", "potrf_dag_kernel", "potrs_kernels", "ewald_pair_kernel",
                                "ewald_self_kernel", "lub_kernel", "gemv_kernel", "pcg_kernel",
                                "gemv_split_f64", "gemv_split_f32", "inner_gmres_graph"};
  return (family >= 0 && family < rsd::kProfFamilies) ? names[family] : "";
}
}
