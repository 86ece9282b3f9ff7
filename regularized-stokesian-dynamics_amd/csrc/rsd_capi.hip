// rsd_capi.hip -- the C ABI (include/rsd.h): host-side lattice logic, the
// drop-in replacements of the reference routines, and the device-resident
// step context.  All numerical work runs in the HIP kernels of rsd_ewald,
// rsd_lub, rsd_dense and rsd_step; this file only orchestrates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rsd_ewald.h"
#include "rsd_free.h"
#include "rsd_gnn.h"
#include "rsd_iter.h"
#include "rsd_pcg.h"
#include "rsd_prof.h"
#include "rsd_internal.h"
#include "rsd_lub.h"
#include "rsd_lubmath.h"
#include "rsd_step.h"

namespace rsd {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_err = buf;
  return RSD_ERR_HIP;
}

// ----------------------------------------------------------- lattice (host)
// tensors.f90:704-726 MAT_INV3
static void mat_inv3(const double* LR, double* LI, double* det) {
  auto L = [&](int i, int j) { return LR[(i - 1) + 3 * (j - 1)]; };
  double d = L(1, 1) * L(2, 2) * L(3, 3) + L(1, 2) * L(2, 3) * L(3, 1) + L(1, 3) * L(2, 1) * L(3, 2) -
             L(1, 3) * L(2, 2) * L(3, 1) - L(2, 3) * L(3, 2) * L(1, 1) - L(3, 3) * L(1, 2) * L(2, 1);
  auto O = [&](int i, int j) -> double& { return LI[(i - 1) + 3 * (j - 1)]; };
  O(1, 1) = (1.0 / d) * (L(2, 2) * L(3, 3) - L(2, 3) * L(3, 2));
  O(1, 2) = (1.0 / d) * (L(1, 3) * L(3, 2) - L(1, 2) * L(3, 3));
  O(1, 3) = (1.0 / d) * (L(1, 2) * L(2, 3) - L(1, 3) * L(2, 2));
  O(2, 1) = (1.0 / d) * (L(2, 3) * L(3, 1) - L(2, 1) * L(3, 3));
  O(2, 2) = (1.0 / d) * (L(1, 1) * L(3, 3) - L(1, 3) * L(3, 1));
  O(2, 3) = (1.0 / d) * (L(1, 3) * L(2, 1) - L(1, 1) * L(2, 3));
  O(3, 1) = (1.0 / d) * (L(2, 1) * L(3, 2) - L(2, 2) * L(3, 1));
  O(3, 2) = (1.0 / d) * (L(1, 2) * L(3, 1) - L(1, 1) * L(3, 2));
  O(3, 3) = (1.0 / d) * (L(1, 1) * L(2, 2) - L(1, 2) * L(2, 1));
  *det = d;
}

// eigenvalues of a symmetric 3x3 (cyclic Jacobi); only extremes are used
static void sym3_eigvals(const double* A, double* ev) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = A[i + 3 * j];
  for (int sweep = 0; sweep < 64; ++sweep) {
    if (fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]) == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          double x = a[k][p], y = a[k][q];
          a[k][p] = c * x - s * y;
          a[k][q] = s * x + c * y;
        }
        for (int k = 0; k < 3; ++k) {
          double x = a[p][k], y = a[q][k];
          a[p][k] = c * x - s * y;
          a[q][k] = s * x + c * y;
        }
      }
  }
  ev[0] = a[0][0];
  ev[1] = a[1][1];
  ev[2] = a[2][2];
}

static void host_set_lattice_skew(double t, const double* lb, double gamma, rsd_lattice* lat) {
  memset(lat, 0, sizeof *lat);
  lat->lr[0] = lb[0];
  lat->lr[4] = lb[1];
  lat->lr[8] = lb[2];
  if (gamma != 0.0) {
    double tshear = lb[0] / (lb[2] * gamma);
    lat->lr[6] = lb[2] * fmod(t, tshear) * gamma;
  }
  double det;
  mat_inv3(lat->lr, lat->li, &det);
  lat->lv = fabs(det);
}

static void host_calc_lattice_inverse(rsd_lattice* lat, double ews_err, int n, const double* radii) {
  lat->sig = pow(lat->lv, 1.0 / 3.0) / sqrt(2.0 * kPi);
  lat->lammda = sqrt(kPi) / pow(lat->lv, 1.0 / 3.0);
  lat->logerr = -log(ews_err);
  double MM[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += lat->lr[k + 3 * i] * lat->lr[k + 3 * j];
      MM[i + 3 * j] = s;
    }
  double ev[3];
  sym3_eigvals(MM, ev);
  double emin = std::min(ev[0], std::min(ev[1], ev[2])), emax = std::max(ev[0], std::max(ev[1], ev[2]));
  lat->nr = (int)(sqrt(2.0 * lat->logerr * lat->sig * lat->sig / emin) + 1.0);
  lat->ni = (int)(sqrt(lat->logerr / (2.0 * kPi * kPi * lat->sig * lat->sig * (1.0 / emax))) + 1.0);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += radii[i] * radii[i] * radii[i];
  lat->phi = s * 4.0 / 3.0 * kPi / lat->lv;
}

static EwaldParams ewald_params(const rsd_lattice& lat, int N, int ld, int corr, double mu_f) {
  EwaldParams p{};
  memcpy(p.LR, lat.lr, sizeof p.LR);
  memcpy(p.LI, lat.li, sizeof p.LI);
  p.LV = lat.lv;
  p.lam = lat.lammda;
  p.phix = 1.0 - 0.2 * lat.phi;
  p.inv_pimu = 1.0 / (kPi * mu_f);
  p.NR = lat.nr;
  p.NI = lat.ni;
  p.N = N;
  p.ld = ld;
  p.correction_method = corr;
  p.offT = 0;
  p.offR = 3 * N;
  p.offQ = 6 * N;
  p.kmax2 = 0;
  return p;
}

// Block columns (of 64 rows) up to which the Schur route forms inv(L_S) inside
// the factorisation's launch instead of by a triangular inverse after it.
constexpr int kFusedInverseMaxTiles = 176;

// ------------------------------------------- allocation-free assembly scratch
// rsd_per_grpery_inv_fri_noalloc keeps its scratch here: 256 bytes for the
// overlap counter, then the k-space weights of up to (2*7+1)^3 wave vectors
// (kKW doubles each), 1.65 MiB per device, held until the process ends.
// Host code never uses the array's name as an address: each device's copy is
// resolved once with hipGetSymbolAddress (with that device current) and the
// resolved pointer is what the kernels receive.
constexpr int kNoAllocMaxNI = 7;
constexpr size_t kNoAllocHead = 32;  // doubles before the k-space weights (the counter lives in the first)
constexpr size_t kNoAllocDoubles =
    kNoAllocHead + (size_t)(2 * kNoAllocMaxNI + 1) * (2 * kNoAllocMaxNI + 1) * (2 * kNoAllocMaxNI + 1) * rsdm::kKW;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
static double* g_scratch_ptr[64];  // per device: the device address of its copy of scratch
static std::mutex g_scratch_mu;

// Resolves the calling thread's current device's copy of scratch (once).
static int noalloc_scratch(double** out) {
  int dev = 0;
  RSD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device index beyond the 64 the library tracks");
    return RSD_ERR_CAPACITY;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    RSD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return RSD_OK;
}

__global__ void zero_int_kernel(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// ----------------------------------------------------------- global state
struct DeviceTables {
  rsdm::LubTables* d = nullptr;
};
static std::mutex g_mu;
static bool g_tables_loaded = false;
static rsdm::LubTables g_tables;
static DeviceTables g_dev_tables[64];

static int current_device(int* dev) {
  RSD_HIP(hipGetDevice(dev));
  return RSD_OK;
}

static int device_tables(rsdm::LubTables** out) {
  if (!g_tables_loaded) {
    set_error("lubrication tables not loaded: call rsd_lub_tables_load first");
    return RSD_ERR_TABLES;
  }
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceTables& t = g_dev_tables[dev & 63];
  if (!t.d) {
    RSD_HIP(hipMalloc(&t.d, sizeof(rsdm::LubTables)));
    RSD_HIP(hipMemcpy(t.d, &g_tables, sizeof(rsdm::LubTables), hipMemcpyHostToDevice));
  }
  *out = t.d;
  return RSD_OK;
}

static int read_table(const char* dir, const char* name, int rows, int cols, double* out) {
  std::string path = std::string(dir) + "/" + name;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) {
    set_error("cannot open " + path);
    return RSD_ERR_TABLES;
  }
  char line[1024];
  if (!fgets(line, sizeof line, f)) {
    fclose(f);
    set_error("empty table " + path);
    return RSD_ERR_TABLES;
  }
  for (int i = 0; i < rows * cols; ++i)
    if (fscanf(f, "%lf", &out[i]) != 1) {
      fclose(f);
      set_error("short table " + path);
      return RSD_ERR_TABLES;
    }
  fclose(f);
  return RSD_OK;
}

// RAII device buffer
struct DBuf {
  double* p = nullptr;
  size_t n = 0;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (count <= n) return RSD_OK;
    if (p) RSD_HIP(hipFree(p));
    p = nullptr;
    RSD_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)));
    n = count;
    return RSD_OK;
  }
};

// host <-> device copies of a column-major sub-block
static int d2h_block(const double* dsrc, int ldd, double* h, int ldh, int m, int n, hipStream_t s) {
  if (m <= 0 || n <= 0) return RSD_OK;
  RSD_HIP(hipMemcpy2DAsync(h, (size_t)ldh * sizeof(double), dsrc, (size_t)ldd * sizeof(double), (size_t)m * sizeof(double),
                           n, hipMemcpyDeviceToHost, s));
  return RSD_OK;
}
static int h2d_block(const double* h, int ldh, double* ddst, int ldd, int m, int n, hipStream_t s) {
  if (m <= 0 || n <= 0) return RSD_OK;
  RSD_HIP(hipMemcpy2DAsync(ddst, (size_t)ldd * sizeof(double), h, (size_t)ldh * sizeof(double), (size_t)m * sizeof(double),
                           n, hipMemcpyHostToDevice, s));
  return RSD_OK;
}

// SPD inverse with upper-triangle input semantics (MATREV) on device data:
// src(n x n, ld lds) -> dst(n x n, ld ldd).  Returns INFO >= 0 or an error.
static int matrev_device(DenseWork& w, const double* src, int n, int lds, double* dst, int ldd, hipStream_t s) {
  const int np = round_up(std::max(n, 1), kTile);
  DBuf A, Ai;
  RSD_TRY(A.alloc((size_t)np * np));
  RSD_TRY(Ai.alloc((size_t)np * np));
  RSD_TRY(dense_reserve(w, np, s));
  RSD_TRY(load_sym_from_upper(src, n, lds, A.p, np, np, s));
  RSD_TRY(dense_info_reset(w, s));
  RSD_TRY(spd_inverse(w, A.p, Ai.p, np, np, s));
  RSD_TRY(copy_block(Ai.p, np, dst, ldd, n, n, s));
  int info = 0;
  RSD_HIP(hipMemcpyAsync(&info, w.d_info, sizeof(int), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipStreamSynchronize(s));
  int werr = 0;
  RSD_TRY(dense_sync_error(w, s, &werr));
  if (werr) {
    set_error("MATREV: a factorisation workgroup ran out of its bounded wait");
    return RSD_ERR_HIP;
  }
  return info == 0x7fffffff ? 0 : info;
}

}  // namespace rsd

using namespace rsd;

// ===================================================================== C ABI
extern "C" {

const char* rsd_last_error(void) { return g_err.c_str(); }
int rsd_version(void) { return 1; }

void rsd_step_config_default(rsd_step_config* cfg, double box_edge) {
  memset(cfg, 0, sizeof *cfg);
  cfg->lb[0] = cfg->lb[1] = cfg->lb[2] = box_edge;
  cfg->gamma = 1.0;
  cfg->ews_err = 1e-4;
  cfg->mu_f = 1.0 / kPi;
  cfg->correction_method = 2;
  cfg->nb = 0;
  cfg->gravity = 0.0;
  cfg->dt_dem0 = 1.0;
  cfg->radii_test = 1.0;
}

int rsd_device_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_error("no HIP device available");
    return RSD_ERR_NODEVICE;
  }
  if (device < 0 || device >= n) {
    set_error("device index out of range");
    return RSD_ERR_ARG;
  }
  RSD_HIP(hipSetDevice(device));
  double* sp = nullptr;  // this device's copy of the allocation-free scratch, resolved at set-up
  return noalloc_scratch(&sp);
}

int rsd_lub_tables_load(const char* dir) {
  if (!dir) return RSD_ERR_ARG;
  rsdm::LubTables t;
  RSD_TRY(read_table(dir, "r2babc.dat", rsdm::kLubAbcRows, 11, t.abc));
  RSD_TRY(read_table(dir, "r2bgh.dat", rsdm::kLubGhRows, 7, t.gh));
  RSD_TRY(read_table(dir, "r2bm.dat", rsdm::kLubMRows, 4, t.m));
  std::lock_guard<std::mutex> lk(g_mu);
  g_tables = t;
  g_tables_loaded = true;
  for (auto& d : g_dev_tables)
    if (d.d) {
      (void)hipFree(d.d);
      d.d = nullptr;
    }
  return RSD_OK;
}

int rsd_set_lattice_skew(double t, const double lb[3], double gamma, rsd_lattice* lat) {
  if (!lb || !lat) return RSD_ERR_ARG;
  host_set_lattice_skew(t, lb, gamma, lat);
  return RSD_OK;
}

int rsd_calc_lattice_inverse(rsd_lattice* lat, double ews_err, int n, const double* radii) {
  if (!lat || n < 0 || (n > 0 && !radii) || !(ews_err > 0.0)) return RSD_ERR_ARG;
  host_calc_lattice_inverse(lat, ews_err, n, radii);
  return RSD_OK;
}

int rsd_per_skew_correction(const rsd_lattice* lat, double r[3], double* dmin) {
  if (!lat || !r || !dmin) return RSD_ERR_ARG;
  double X[3], RC[3];
  for (int i = 0; i < 3; ++i) {
    X[i] = lat->li[i] * r[0] + lat->li[i + 3] * r[1] + lat->li[i + 6] * r[2];
    X[i] = fmod(X[i], 1.0);
    if (X[i] < 0.0) X[i] += 1.0;
  }
  for (int i = 0; i < 3; ++i) RC[i] = lat->lr[i] * X[0] + lat->lr[i + 3] * X[1] + lat->lr[i + 6] * X[2];
  double D = RC[0] * RC[0] + RC[1] * RC[1] + RC[2] * RC[2], best[3] = {RC[0], RC[1], RC[2]};
  for (int m = 1; m < 8; ++m) {
    int m1 = (m >> 2) & 1, m2 = (m >> 1) & 1, m3 = m & 1;
    double d[3], dd = 0.0;
    for (int i = 0; i < 3; ++i) {
      d[i] = RC[i] - (lat->lr[i] * m1 + lat->lr[i + 3] * m2 + lat->lr[i + 6] * m3);
      dd += d[i] * d[i];
    }
    if (dd < D) {
      D = dd;
      memcpy(best, d, sizeof best);
    }
  }
  memcpy(r, best, sizeof best);
  *dmin = sqrt(D);
  return RSD_OK;
}

}  // extern "C"

// PER_GRPERY_INV_FRI on host arrays; nr >= 0 replaces the rule's NR and NI,
// kmax2 > 0 cuts the reciprocal cube to |n|^2 <= kmax2
static int per_grpery_impl(int n, const double* pos, const double* radii, rsd_lattice* lat, double ews_err, int nr,
                           int ni, int kmax2, int correction_method, double mu_f, double* app, double* apq,
                           double* aqq, int* n_overlap) {
  if (n < 1 || !pos || !radii || !lat || !app || !apq || !aqq || !(ews_err > 0.0)) return RSD_ERR_ARG;
  host_calc_lattice_inverse(lat, ews_err, n, radii);
  if (nr >= 0) {
    lat->nr = nr;
    lat->ni = ni;
  }
  const int n11 = 11 * n, n6 = 6 * n, n5 = 5 * n;
  hipStream_t s = nullptr;
  DBuf dpos, drad, G, kw;
  int* dov = nullptr;
  RSD_TRY(dpos.alloc(3 * (size_t)n));
  RSD_TRY(drad.alloc(n));
  RSD_TRY(G.alloc((size_t)n11 * n11));
  const int side = 2 * lat->ni + 1;
  RSD_TRY(kw.alloc((size_t)side * side * side * rsdm::kKW));
  RSD_HIP(hipMalloc(&dov, sizeof(int)));
  RSD_HIP(hipMemsetAsync(dov, 0, sizeof(int), s));
  RSD_HIP(hipMemcpyAsync(dpos.p, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_HIP(hipMemcpyAsync(drad.p, radii, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  EwaldParams p = ewald_params(*lat, n, n11, correction_method, mu_f);
  p.kmax2 = kmax2;
  int rc = ewald_assemble(p, dpos.p, drad.p, kw.p, G.p, dov, s);
  if (!rc) rc = d2h_block(G.p, n11, app, n6, n6, n6, s);
  if (!rc) rc = d2h_block(G.p + (size_t)n6 * n11, n11, apq, n6, n6, n5, s);
  if (!rc) rc = d2h_block(G.p + n6 + (size_t)n6 * n11, n11, aqq, n5, n5, n5, s);
  int nov = 0;
  if (!rc) {
    hipError_t e = hipMemcpyAsync(&nov, dov, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "per_grpery_inv_fri sync", __FILE__, __LINE__);
  }
  (void)hipFree(dov);
  if (n_overlap) *n_overlap = nov;
  return rc;
}

extern "C" {

int rsd_per_grpery_inv_fri(int n, const double* pos, const double* radii, rsd_lattice* lat, double ews_err,
                           int correction_method, double mu_f, double* app, double* apq, double* aqq,
                           int* n_overlap) {
  return per_grpery_impl(n, pos, radii, lat, ews_err, -1, -1, 0, correction_method, mu_f, app, apq, aqq, n_overlap);
}

int rsd_per_grpery_inv_fri_noalloc(int n, const double* d_pos, const double* d_radii, const rsd_lattice* lat,
                                   int correction_method, double mu_f, double* d_grmobmx, int ld, int* d_n_overlap,
                                   void* stream) {
  if (n < 1 || !d_pos || !d_radii || !lat || !d_grmobmx || ld < 11 * n) return RSD_ERR_ARG;
  if (!(lat->lv > 0.0) || !(lat->lammda > 0.0) || lat->nr < 0 || lat->ni < 0 || !(mu_f > 0.0)) {
    set_error("rsd_per_grpery_inv_fri_noalloc: lat must hold the Ewald constants (rsd_calc_lattice_inverse)");
    return RSD_ERR_ARG;
  }
  if (lat->ni > kNoAllocMaxNI) {
    set_error("rsd_per_grpery_inv_fri_noalloc: NI = " + std::to_string(lat->ni) + " needs more k-space scratch than " +
              "the " + std::to_string(kNoAllocMaxNI) + " the allocation-free entry point holds; raise ews_err or use " +
              "rsd_per_grpery_inv_fri");
    return RSD_ERR_CAPACITY;
  }
  double* sp = nullptr;
  RSD_TRY(noalloc_scratch(&sp));
  hipStream_t s = static_cast<hipStream_t>(stream);
  int* counter = d_n_overlap ? d_n_overlap : reinterpret_cast<int*>(sp);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, s, counter);
  RSD_HIP(hipGetLastError());
  EwaldParams p = ewald_params(*lat, n, ld, correction_method, mu_f);
  return ewald_assemble(p, d_pos, d_radii, sp + kNoAllocHead, d_grmobmx, counter, s, /*profiled=*/false);
}

int rsd_per_grpery_inv_fri_images(int n, const double* pos, const double* radii, rsd_lattice* lat, double ews_err,
                                  int nr, int ni, int kmax2, int correction_method, double mu_f, double* app,
                                  double* apq, double* aqq, int* n_overlap) {
  if (nr < 0 || ni < 0 || nr > 64 || ni > 64 || kmax2 < 0) return RSD_ERR_ARG;
  return per_grpery_impl(n, pos, radii, lat, ews_err, nr, ni, kmax2, correction_method, mu_f, app, apq, aqq,
                         n_overlap);
}

int rsd_grpery_inv_fri(int n, const double* pos, const double* radii, int correction_method, double mu_f,
                       double* app, double* apq, double* aqq, int* n_overlap) {
  if (n < 1 || !pos || !radii || !app || !apq || !aqq || !(mu_f > 0.0)) return RSD_ERR_ARG;
  const int n11 = 11 * n, n6 = 6 * n, n5 = 5 * n;
  hipStream_t s = nullptr;
  DBuf dpos, drad, G;
  int* dov = nullptr;
  RSD_TRY(dpos.alloc(3 * (size_t)n));
  RSD_TRY(drad.alloc(n));
  RSD_TRY(G.alloc((size_t)n11 * n11));
  RSD_HIP(hipMalloc(&dov, sizeof(int)));
  RSD_HIP(hipMemsetAsync(dov, 0, sizeof(int), s));
  RSD_HIP(hipMemcpyAsync(dpos.p, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_HIP(hipMemcpyAsync(drad.p, radii, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  FreeParams fp{1.0 / (kPi * mu_f), n, n11, correction_method, 0, 3 * n, 6 * n};
  int rc = free_assemble(fp, dpos.p, drad.p, G.p, dov, s);
  if (!rc) rc = d2h_block(G.p, n11, app, n6, n6, n6, s);
  if (!rc) rc = d2h_block(G.p + (size_t)n6 * n11, n11, apq, n6, n6, n5, s);
  if (!rc) rc = d2h_block(G.p + n6 + (size_t)n6 * n11, n11, aqq, n5, n5, n5, s);
  int nov = 0;
  if (!rc) {
    hipError_t e = hipMemcpyAsync(&nov, dov, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "grpery_inv_fri sync", __FILE__, __LINE__);
  }
  (void)hipFree(dov);
  if (n_overlap) *n_overlap = nov;
  return rc;
}

int rsd_wall_grpery_inv_fri(int n, const double* pos, const double* radii, double mu_f, double* tpp) {
  if (n < 1 || !pos || !radii || !tpp || !(mu_f > 0.0)) return RSD_ERR_ARG;
  const int n3 = 3 * n;
  hipStream_t s = nullptr;
  DBuf dpos, drad, T;
  RSD_TRY(dpos.alloc(3 * (size_t)n));
  RSD_TRY(drad.alloc(n));
  RSD_TRY(T.alloc((size_t)n3 * n3));
  RSD_HIP(hipMemcpyAsync(dpos.p, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_HIP(hipMemcpyAsync(drad.p, radii, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_TRY(wall_tt(n, mu_f, dpos.p, drad.p, T.p, n3, false, s));
  RSD_TRY(d2h_block(T.p, n3, tpp, n3, n3, n3, s));
  RSD_HIP(hipStreamSynchronize(s));
  return RSD_OK;
}

int rsd_lubmxcalc(int n, int nb, int is_period, const double* pos, const double* radii, const rsd_lattice* lat,
                  double mu_f, double* rfu, double* rfe, double* rse, int* n_contact) {
  if (n < 1 || nb < 0 || nb > n || !pos || !radii || !lat || !rfu || !rfe || !rse) return RSD_ERR_ARG;
  rsdm::LubTables* tab = nullptr;
  RSD_TRY(device_tables(&tab));
  const int n6 = 6 * n, n5 = 5 * n;
  hipStream_t s = nullptr;
  DBuf dpos, drad, dpp, dpq, dqq;
  RSD_TRY(dpos.alloc(3 * (size_t)n));
  RSD_TRY(drad.alloc(n));
  RSD_TRY(dpp.alloc((size_t)n6 * n6));
  RSD_TRY(dpq.alloc((size_t)n6 * n5));
  RSD_TRY(dqq.alloc((size_t)n5 * n5));
  int* dct = nullptr;
  RSD_HIP(hipMalloc(&dct, sizeof(int)));
  RSD_HIP(hipMemsetAsync(dct, 0, sizeof(int), s));
  RSD_HIP(hipMemcpyAsync(dpos.p, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_HIP(hipMemcpyAsync(drad.p, radii, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  int rc = launch_fill(dpp.p, (size_t)n6 * n6, 0.0, s);
  if (!rc) rc = launch_fill(dpq.p, (size_t)n6 * n5, 0.0, s);
  if (!rc) rc = launch_fill(dqq.p, (size_t)n5 * n5, 0.0, s);
  LubParams lp{};
  memcpy(lp.LR, lat->lr, sizeof lp.LR);
  memcpy(lp.LI, lat->li, sizeof lp.LI);
  lp.scale = 6.0 * (kPi * mu_f);
  lp.N = n;
  lp.Nb = nb;
  lp.is_period = is_period;
  lp.PP = dpp.p;
  lp.ldPP = n6;
  lp.PQ = dpq.p;
  lp.ldPQ = n6;
  lp.QQ = dqq.p;
  lp.ldQQ = n5;
  lp.pq_zero = 1;  // all three targets were zero-filled above
  if (!rc) rc = lub_add(lp, dpos.p, drad.p, tab, dct, s);
  if (!rc) rc = d2h_block(dpp.p, n6, rfu, n6, n6, n6, s);
  if (!rc) rc = d2h_block(dpq.p, n6, rfe, n6, n6, n5, s);
  if (!rc) rc = d2h_block(dqq.p, n5, rse, n5, n5, n5, s);
  int nc = 0;
  if (!rc) {
    hipError_t e = hipMemcpyAsync(&nc, dct, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "lubmxcalc sync", __FILE__, __LINE__);
  }
  (void)hipFree(dct);
  if (n_contact) *n_contact = nc;
  return rc;
}

static DenseWork& host_api_dense() {
  static thread_local DenseWork w;
  return w;
}

int rsd_matrev(int n, double* a) {
  if (n < 1 || !a) return RSD_ERR_ARG;
  hipStream_t s = nullptr;
  DBuf A;
  RSD_TRY(A.alloc((size_t)n * n));
  RSD_TRY(h2d_block(a, n, A.p, n, n, n, s));
  int info = matrev_device(host_api_dense(), A.p, n, n, A.p, n, s);
  if (info < 0) return info;
  RSD_TRY(d2h_block(A.p, n, a, n, n, n, s));
  RSD_HIP(hipStreamSynchronize(s));
  return info;
}

int rsd_cholesky(int n, double* a) {
  if (n < 1 || !a) return RSD_ERR_ARG;
  const int np = round_up(n, kTile);
  hipStream_t s = nullptr;
  DenseWork& w = host_api_dense();
  DBuf A;
  RSD_TRY(A.alloc((size_t)np * np));
  RSD_TRY(h2d_block(a, n, A.p, np, n, n, s));
  RSD_TRY(set_identity_padding(A.p, n, np, np, s));
  RSD_TRY(dense_reserve(w, np, s));
  RSD_TRY(dense_info_reset(w, s));
  RSD_TRY(potrf_lower(w, A.p, np, np, s));
  RSD_TRY(zero_strict_upper(A.p, np, np, s));
  int info = 0;
  RSD_HIP(hipMemcpyAsync(&info, w.d_info, sizeof(int), hipMemcpyDeviceToHost, s));
  RSD_TRY(d2h_block(A.p, np, a, n, n, n, s));
  RSD_HIP(hipStreamSynchronize(s));
  int werr = 0;
  RSD_TRY(dense_sync_error(w, s, &werr));
  if (werr) {
    set_error("CHOLESKY: a factorisation workgroup ran out of its bounded wait");
    return RSD_ERR_HIP;
  }
  return info == 0x7fffffff ? 0 : info;
}

int rsd_mob_to_resist_fts(int n, const double* app, const double* apq, const double* aqq, double* rpp, double* rpq,
                          double* rqq) {
  if (n < 1 || !app || !apq || !aqq || !rpp || !rpq || !rqq) return RSD_ERR_ARG;
  const int n11 = 11 * n, n6 = 6 * n, n5 = 5 * n;
  hipStream_t s = nullptr;
  DBuf M;
  RSD_TRY(M.alloc((size_t)n11 * n11));
  // matrices.f90:400-403: [APP APQ; APQ^T AQQ]; only the upper triangle is read
  RSD_TRY(h2d_block(app, n6, M.p, n11, n6, n6, s));
  RSD_TRY(h2d_block(apq, n6, M.p + (size_t)n6 * n11, n11, n6, n5, s));
  RSD_TRY(h2d_block(aqq, n5, M.p + n6 + (size_t)n6 * n11, n11, n5, n5, s));
  int info = matrev_device(host_api_dense(), M.p, n11, n11, M.p, n11, s);
  if (info < 0) return info;
  RSD_TRY(d2h_block(M.p, n11, rpp, n6, n6, n6, s));
  RSD_TRY(d2h_block(M.p + (size_t)n6 * n11, n11, rpq, n6, n6, n5, s));
  RSD_TRY(d2h_block(M.p + n6 + (size_t)n6 * n11, n11, rqq, n5, n5, n5, s));
  RSD_HIP(hipStreamSynchronize(s));
  return info;
}

// mu = RPQ^T inv(RPP) RPQ - RQQ on device buffers (leading dims n6 / n5)
static int arrangement_device(DenseWork& w, const double* dRPP, const double* dRPQ, const double* dRQQ, int n,
                              double* dMU, hipStream_t s) {
  const int n6 = 6 * n, n5 = 5 * n;
  const int n6p = round_up(n6, kTile), n5p = round_up(n5, kTile);
  DBuf Pi, Qp, T, Mu;
  RSD_TRY(Pi.alloc((size_t)n6p * n6p));
  RSD_TRY(Qp.alloc((size_t)n6p * n5p));
  RSD_TRY(T.alloc((size_t)n6p * n5p));
  RSD_TRY(Mu.alloc((size_t)n5p * n5p));
  int info = matrev_device(w, dRPP, n6, n6, Pi.p, n6p, s);
  if (info < 0) return info;
  RSD_TRY(set_identity_padding(Pi.p, n6, n6p, n6p, s));
  RSD_TRY(launch_fill(Qp.p, (size_t)n6p * n5p, 0.0, s));
  RSD_TRY(copy_block(dRPQ, n6, Qp.p, n6p, n6, n5, s));
  GemmDesc g{};
  g.A = Pi.p;
  g.B = Qp.p;
  g.C = T.p;
  g.lda = n6p;
  g.ldb = n6p;
  g.ldc = n6p;
  g.mt = n6p / kTile;
  g.nt = n5p / kTile;
  g.kt = n6p / kTile;
  g.alpha = 1.0;
  g.beta = 0.0;
  RSD_TRY(gemm_batched(w, {g}, false, false, 0, s));
  // Mu = Qp^T T, then subtract RQQ
  GemmDesc h{};
  h.A = Qp.p;
  h.B = T.p;
  h.C = Mu.p;
  h.lda = n6p;
  h.ldb = n6p;
  h.ldc = n5p;
  h.mt = n5p / kTile;
  h.nt = n5p / kTile;
  h.kt = n6p / kTile;
  h.alpha = 1.0;
  h.beta = 0.0;
  RSD_TRY(gemm_batched(w, {h}, true, false, 0, s));
  RSD_TRY(copy_block(Mu.p, n5p, dMU, n5, n5, n5, s));
  RSD_TRY(matrix_axpy(-1.0, dRQQ, n5, dMU, n5, n5, n5, s));
  return info;
}

int rsd_arrangement(int n, const double* rpp, const double* rpq, const double* rqq, double* mu) {
  if (n < 1 || !rpp || !rpq || !rqq || !mu) return RSD_ERR_ARG;
  const int n6 = 6 * n, n5 = 5 * n;
  hipStream_t s = nullptr;
  DBuf P, Q, R, M;
  RSD_TRY(P.alloc((size_t)n6 * n6));
  RSD_TRY(Q.alloc((size_t)n6 * n5));
  RSD_TRY(R.alloc((size_t)n5 * n5));
  RSD_TRY(M.alloc((size_t)n5 * n5));
  RSD_TRY(h2d_block(rpp, n6, P.p, n6, n6, n6, s));
  RSD_TRY(h2d_block(rpq, n6, Q.p, n6, n6, n5, s));
  RSD_TRY(h2d_block(rqq, n5, R.p, n5, n5, n5, s));
  int info = arrangement_device(host_api_dense(), P.p, Q.p, R.p, n, M.p, s);
  if (info < 0) return info;
  RSD_TRY(d2h_block(M.p, n5, mu, n5, n5, n5, s));
  RSD_HIP(hipStreamSynchronize(s));
  return info;
}

}  // extern "C"

// ============================================== STEPPER_NOGmres (host arrays)
extern "C" int rsd_stepper_nogmres(int n, const double* app, const double* apq, const double* aqq, const double* rfu,
                                   const double* rfe, const double* rse, const double* fe, const double* ein,
                                   double* u_par, double* sijn_hyd, double* rpp) {
  if (n < 1 || !app || !apq || !aqq || !rfu || !rfe || !rse || !fe || !ein || !u_par || !sijn_hyd) return RSD_ERR_ARG;
  const int n11 = 11 * n, n6 = 6 * n, n5 = 5 * n;
  hipStream_t s = nullptr;
  DenseWork& w = host_api_dense();
  DBuf M, RPP, RPQ, RQQ, APPR, MU, vfe, vein, vc, vu, vs;
  RSD_TRY(M.alloc((size_t)n11 * n11));
  RSD_TRY(RPP.alloc((size_t)n6 * n6));
  RSD_TRY(RPQ.alloc((size_t)n6 * n5));
  RSD_TRY(RQQ.alloc((size_t)n5 * n5));
  RSD_TRY(APPR.alloc((size_t)n6 * n6));
  RSD_TRY(MU.alloc((size_t)n5 * n5));
  RSD_TRY(vfe.alloc(n6));
  RSD_TRY(vein.alloc(n5));
  RSD_TRY(vc.alloc(n6));
  RSD_TRY(vu.alloc(n6));
  RSD_TRY(vs.alloc(n5));
  // MobToResist_FTS (matrices.f90:393-416)
  RSD_TRY(h2d_block(app, n6, M.p, n11, n6, n6, s));
  RSD_TRY(h2d_block(apq, n6, M.p + (size_t)n6 * n11, n11, n6, n5, s));
  RSD_TRY(h2d_block(aqq, n5, M.p + n6 + (size_t)n6 * n11, n11, n5, n5, s));
  int info = matrev_device(w, M.p, n11, n11, M.p, n11, s);
  if (info < 0) return info;
  // arrangResist (matrices.f90:218-228)
  RSD_TRY(copy_block(M.p, n11, RPP.p, n6, n6, n6, s));
  RSD_TRY(copy_block(M.p + (size_t)n6 * n11, n11, RPQ.p, n6, n6, n5, s));
  RSD_TRY(copy_block(M.p + n6 + (size_t)n6 * n11, n11, RQQ.p, n5, n5, n5, s));
  DBuf L;
  RSD_TRY(L.alloc((size_t)n6 * n6));
  RSD_TRY(h2d_block(rfu, n6, L.p, n6, n6, n6, s));
  RSD_TRY(matrix_axpy(1.0, L.p, n6, RPP.p, n6, n6, n6, s));
  RSD_TRY(h2d_block(rfe, n6, L.p, n6, n6, n5, s));
  RSD_TRY(matrix_axpy(1.0, L.p, n6, RPQ.p, n6, n6, n5, s));
  RSD_TRY(h2d_block(rse, n5, L.p, n5, n5, n5, s));
  RSD_TRY(matrix_axpy(1.0, L.p, n5, RQQ.p, n5, n5, n5, s));
  // U_par = inv(RPP) (Fe + RPQ EIN)   (stepper.f90:387-395)
  int info2 = matrev_device(w, RPP.p, n6, n6, APPR.p, n6, s);
  if (info2 < 0) return info2;
  RSD_HIP(hipMemcpyAsync(vfe.p, fe, n6 * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_HIP(hipMemcpyAsync(vein.p, ein, n5 * sizeof(double), hipMemcpyHostToDevice, s));
  RSD_HIP(hipMemcpyAsync(vc.p, vfe.p, n6 * sizeof(double), hipMemcpyDeviceToDevice, s));
  RSD_TRY(gemv(w, RPQ.p, n6, n6, n5, false, 1.0, vein.p, 1.0, vc.p, s));
  RSD_TRY(gemv(w, APPR.p, n6, n6, n6, false, 1.0, vc.p, 0.0, vu.p, s));
  // SijN_hyd = mu EIN, mu = RPQ^T inv(RPP) RPQ - RQQ   (stepper.f90:401-403)
  int info3 = arrangement_device(w, RPP.p, RPQ.p, RQQ.p, n, MU.p, s);
  if (info3 < 0) return info3;
  RSD_TRY(gemv(w, MU.p, n5, n5, n5, false, 1.0, vein.p, 0.0, vs.p, s));
  RSD_HIP(hipMemcpyAsync(u_par, vu.p, n6 * sizeof(double), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipMemcpyAsync(sijn_hyd, vs.p, n5 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (rpp) RSD_TRY(d2h_block(RPP.p, n6, rpp, n6, n6, n6, s));
  RSD_HIP(hipStreamSynchronize(s));
  return info ? info : (info2 ? info2 : info3);
}

// ===================================================== device-resident step
struct rsd_ctx {
  int dev = 0;
  hipStream_t s = nullptr;
  hipStream_t s2 = nullptr;  // side stream: work that overlaps the main stream's (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_copied = nullptr;
  int N = 0, n11 = 0, n11p = 0, n6 = 0, n6p = 0, n5 = 0;
  rsd_step_config cfg{};
  std::vector<double> h_radii;
  DenseWork dw;
  DBuf pos, radii, upos, upar, shyd, ein, fe, c, y, F, UU, yeta;
  DBuf G, W, RPP, RPQ, RQQ, kw;
  DBuf bvec;             // sub-step right-hand sides for the CG solves (rows padded with zeros)
  DBuf pos_save, upos_save;  // the state at the first sub-step, for a direct redo
  DBuf xhist;            // the solutions of the sub-step before the last (CG initial guesses)
  // Schur-complement route: W_PP = inv(A - B inv(C) B^T) for G = [A B; B^T C],
  // from the factor of [C B^T; B A] (held in W's storage)
  int resist = RSD_RESIST_SCHUR;
  // Brownian motion (stepper.f90:72-86): kBT > 0 turns it on; z from the
  // host for the next step when given, else from the counter-based generator
  double kbt = 0.0;
  unsigned long long brown_seed = 0, brown_step = 0;
  std::vector<double> z_next;  // the next step's z from the host (empty: draw it)
  DBuf bz, ub, sb, A3, Tb, Gb, bpos1, bpos2;  // z (3N), UB (6N, padded), SijN_B (5N), scratch
  int* brown_info = nullptr;   // INFO of the last Brownian CHOLESKY (0x7fffffff: none failed)
  int* g_info = nullptr;       // INFO of the grand mobility's factorisation (0x7fffffff: factored)
  int img_nr = -1, img_ni = -1, img_kmax2 = 0;  // explicit Ewald image sets (img_nr < 0: the reference rule)
  // iterative route (RSD_RESIST_ITERATIVE): GMRES on the mobility form
  IterWork iw;
  int it_precond = kIterMixed, it_max = 400, it_inner = 20, it_restart = 30;
  bool it_graphs = true;  // replay the mixed preconditioners' inner GMRES as a captured HIP graph
  double it_tol = 1e-12;
  int last_it_iters = 0, last_it_fallback = 0;
  double last_it_relres = 0.0;
  static constexpr int kBsrCap = 96;  // lubrication neighbours per particle the block-sparse rows hold
  DBuf bsr_blk, bsr_self, xit, bit, wit;
  int* bsr_ix = nullptr;   // [0, N) counts, [N, N + N cap) columns, then the overflow word
  float* bsr32 = nullptr;  // FP32 copies: blocks, self blocks, then M
  size_t bsr32_cap = 0;
  GnnWork gw;                 // the learned preconditioner's buffers
  rsd_gnn* gnn = nullptr;     // the attached model (rsd_ctx_set_gnn), not owned
  bool iter_ready = false;    // rsd_ctx_prepare_iterative has assembled the system
  int is_period = 1;     // IsPeriod/simplePeriod: Ewald under Lees-Edwards images, or free space
  int wall = 0;          // Wall_method: the wall images' TT mobility added to APP
  int n5p = 0, npq = 0;  // npq = n5p + n6p, the order of [C B^T; B A]
  DBuf WPP;              // W_PP, n6p x n6p, full, identity padding
  DBuf dinvC;            // the 64x64 diagonal inverses of L_C
  DBuf wpq_ein, ec, v5, v6;  // W_PQ EIN, inv(L_C) EIN, work vectors
  DBuf lpe, lqe;             // the lubrication's RPQ EIN (6N) and RQQ EIN (5N), formed by lub_add
  PcgWork pw;
  int solver = RSD_SUBSTEP_PCG;
  double cg_tol = 1e-14;
  int cg_max_iter = 60;
  int last_cg_iters = 0, last_fallbacks = 0;
  int last_redo = 0;     // 1 when the last step's Schur factorisation failed and the step was redone explicitly
  int* flags = nullptr;  // [0] overlaps, [1] contacts, [2] collision
  int* h_flags = nullptr;  // pinned copy of flags
  struct StepEnd {        // what the host reads at the end of a step, in one wait
    double yeta[5];
    int info;
    int wait_err[2];      // bounded-wait error words: triangular solves, factorisation
    unsigned cg_err;      // bounded-wait error word of the one-launch CG
    int brown_info;       // INFO of the Brownian CHOLESKY (0x7fffffff: none failed)
  };
  StepEnd* h_end = nullptr;  // pinned, written by step_end_kernel
  StepEnd* d_end = nullptr;  // its device address
  std::vector<int> cg_pred;  // CG iterations each status slot took last time
  hipEvent_t ev[8]{};
  double stage_ms[6]{};
  int last_nt = 1;
  bool has_state = false;
};

extern "C" int rsd_ctx_create(rsd_ctx** out, int n, const rsd_step_config* cfg) {
  if (!out || n < 1 || !cfg) return RSD_ERR_ARG;
  if (cfg->nb < 0 || cfg->nb > n || !(cfg->ews_err > 0.0) || !(cfg->mu_f > 0.0)) return RSD_ERR_ARG;
  std::unique_ptr<rsd_ctx> c(new rsd_ctx);
  RSD_HIP(hipGetDevice(&c->dev));
  RSD_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  RSD_HIP(hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking));
  RSD_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  RSD_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  RSD_HIP(hipEventCreateWithFlags(&c->ev_copied, hipEventDisableTiming));
  c->N = n;
  c->cfg = *cfg;
  c->n11 = 11 * n;
  c->n11p = round_up(c->n11, kTile);
  c->n6 = 6 * n;
  c->n6p = round_up(c->n6, kTile);
  c->n5 = 5 * n;
  RSD_TRY(c->pos.alloc(3 * (size_t)n));
  RSD_TRY(c->radii.alloc(n));
  RSD_TRY(c->upos.alloc(6 * (size_t)n));
  RSD_TRY(c->upar.alloc(2 * (size_t)c->n6p));  // two solve columns, rows padded to n6p
  RSD_TRY(c->shyd.alloc(5 * (size_t)n));
  RSD_TRY(c->ein.alloc(5 * (size_t)n));
  RSD_TRY(c->fe.alloc(6 * (size_t)n));
  RSD_TRY(c->c.alloc(6 * (size_t)n));
  RSD_TRY(c->y.alloc(6 * (size_t)n));
  RSD_TRY(c->F.alloc(6 * (size_t)n));
  RSD_TRY(c->UU.alloc(6 * (size_t)n));
  RSD_TRY(c->yeta.alloc(8));
  c->n5p = round_up(c->n5, kTile);
  c->npq = c->n5p + c->n6p;
  const size_t nbig = (size_t)std::max(c->n11p, c->npq);
  RSD_TRY(c->G.alloc((size_t)c->n11p * c->n11p));
  RSD_TRY(c->W.alloc(nbig * nbig));
  RSD_TRY(c->RPP.alloc((size_t)c->n6p * c->n6p));
  RSD_TRY(c->RPQ.alloc((size_t)c->n6 * c->n5));
  RSD_TRY(c->RQQ.alloc((size_t)c->n5 * c->n5));
  RSD_TRY(dense_reserve(c->dw, (int)nbig, c->s));
  RSD_HIP(hipMalloc(&c->flags, 4 * sizeof(int)));
  RSD_HIP(hipHostMalloc(&c->h_flags, 4 * sizeof(int)));
  RSD_HIP(hipHostMalloc(&c->h_end, sizeof(rsd_ctx::StepEnd), hipHostMallocMapped));
  RSD_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_end), c->h_end, 0));
  c->cg_pred.assign(kPcgMaxSlots, 15);
  for (auto& e : c->ev) RSD_HIP(hipEventCreate(&e));
  RSD_TRY(launch_fill(c->upos.p, 6 * (size_t)n, 0.0, c->s));
  RSD_TRY(launch_fill(c->upar.p, 2 * (size_t)c->n6p, 0.0, c->s));
  RSD_TRY(c->bvec.alloc(2 * (size_t)c->n6p));
  RSD_TRY(c->pos_save.alloc(3 * (size_t)n));
  RSD_TRY(c->xhist.alloc(2 * (size_t)c->n6p));
  RSD_TRY(c->WPP.alloc((size_t)c->n6p * c->n6p));
  RSD_TRY(c->dinvC.alloc((size_t)c->n5p * 64));
  RSD_TRY(c->wpq_ein.alloc(c->n6p));
  RSD_TRY(c->ec.alloc(c->n5p));
  RSD_TRY(c->v5.alloc(c->n5p));
  RSD_TRY(c->v6.alloc(c->n6p));
  RSD_TRY(c->lpe.alloc(c->n6p));
  RSD_TRY(c->lqe.alloc(c->n5p));
  RSD_TRY(c->upos_save.alloc(6 * (size_t)n));
  RSD_TRY(launch_fill(c->bvec.p, 2 * (size_t)c->n6p, 0.0, c->s));
  *out = c.release();
  return RSD_OK;
}

extern "C" int rsd_ctx_destroy(rsd_ctx* c) {
  if (!c) return RSD_OK;
  (void)hipStreamSynchronize(c->s);
  if (c->s2) (void)hipStreamSynchronize(c->s2);
  dense_release(c->dw);
  pcg_release(c->pw);
  if (c->flags) (void)hipFree(c->flags);
  if (c->brown_info) (void)hipFree(c->brown_info);
  if (c->g_info) (void)hipFree(c->g_info);
  if (c->bsr_ix) (void)hipFree(c->bsr_ix);
  if (c->bsr32) (void)hipFree(c->bsr32);
  gnn_work_release(c->gw);
  iter_release(c->iw);
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->h_end) (void)hipHostFree(c->h_end);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_copied) (void)hipEventDestroy(c->ev_copied);
  if (c->s2) (void)hipStreamDestroy(c->s2);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_state(rsd_ctx* c, const double* conf, const double* radii, const double* u_pos) {
  if (!c || !conf || !radii) return RSD_ERR_ARG;
  const int N = c->N;
  for (int i = 0; i < N; ++i)
    if (!(radii[i] > 0.0)) {
      set_error("radii must be positive");
      return RSD_ERR_ARG;
    }
  c->h_radii.assign(radii, radii + N);
  c->iter_ready = false;
  RSD_HIP(hipMemcpyAsync(c->pos.p, conf, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, c->s));
  RSD_HIP(hipMemcpyAsync(c->radii.p, radii, (size_t)N * sizeof(double), hipMemcpyHostToDevice, c->s));
  if (u_pos)
    RSD_HIP(hipMemcpyAsync(c->upos.p, u_pos, 6 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, c->s));
  RSD_HIP(hipStreamSynchronize(c->s));
  c->has_state = true;
  return RSD_OK;
}

extern "C" int rsd_ctx_get_state(rsd_ctx* c, double* conf, double* u_pos) {
  if (!c) return RSD_ERR_ARG;
  if (conf) RSD_HIP(hipMemcpyAsync(conf, c->pos.p, 3 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (u_pos) RSD_HIP(hipMemcpyAsync(u_pos, c->upos.p, 6 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  RSD_HIP(hipStreamSynchronize(c->s));
  return RSD_OK;
}

extern "C" int rsd_ctx_get_last(rsd_ctx* c, double* u_par, double* sijn_hyd, int* nt_dem) {
  if (!c) return RSD_ERR_ARG;
  if (u_par) RSD_HIP(hipMemcpyAsync(u_par, c->upar.p, 6 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (sijn_hyd)
    RSD_HIP(hipMemcpyAsync(sijn_hyd, c->shyd.p, 5 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  RSD_HIP(hipStreamSynchronize(c->s));
  if (nt_dem) *nt_dem = c->last_nt;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_substep_solver(rsd_ctx* c, int mode, double tol, int max_iter) {
  if (!c || (mode != RSD_SUBSTEP_DIRECT && mode != RSD_SUBSTEP_PCG && mode != RSD_SUBSTEP_PCG_ONE_LAUNCH))
    return RSD_ERR_ARG;
  if (mode != RSD_SUBSTEP_DIRECT && (!(tol > 0.0) || !(tol < 1.0) || max_iter < 0)) return RSD_ERR_ARG;
  c->solver = mode;
  if (mode != RSD_SUBSTEP_DIRECT) {
    c->cg_tol = tol;
    c->cg_max_iter = max_iter;
  }
  return RSD_OK;
}

extern "C" int rsd_ctx_set_brownian(rsd_ctx* c, double kbt, unsigned long long seed) {
  if (!c || !(kbt >= 0.0)) return RSD_ERR_ARG;
  c->kbt = kbt;
  c->brown_seed = seed;
  c->brown_step = 0;
  c->z_next.clear();
  return RSD_OK;
}

extern "C" int rsd_ctx_set_brownian_z(rsd_ctx* c, const double* z) {
  if (!c || !z) return RSD_ERR_ARG;
  c->z_next.assign(z, z + 3 * (size_t)c->N);
  return RSD_OK;
}

extern "C" int rsd_ctx_get_brownian(rsd_ctx* c, double* z, double* u_brown, double* sijn_b) {
  if (!c) return RSD_ERR_ARG;
  if (!(c->kbt > 0.0) || !c->bz.p) {
    set_error("rsd_ctx_get_brownian: no Brownian step has run");
    return RSD_ERR_ARG;
  }
  const size_t n3 = 3 * (size_t)c->N;
  if (z) RSD_HIP(hipMemcpyAsync(z, c->bz.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (u_brown) RSD_HIP(hipMemcpyAsync(u_brown, c->ub.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (sijn_b) RSD_HIP(hipMemcpyAsync(sijn_b, c->sb.p, 5 * (size_t)c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  RSD_HIP(hipStreamSynchronize(c->s));
  return RSD_OK;
}

extern "C" int rsd_ctx_set_ewald_images(rsd_ctx* c, int nr, int ni, int kmax2) {
  if (!c || (nr >= 0 && (ni < 0 || ni > 64 || nr > 64)) || kmax2 < 0) return RSD_ERR_ARG;
  c->img_nr = nr < 0 ? -1 : nr;
  c->img_ni = nr < 0 ? -1 : ni;
  c->img_kmax2 = kmax2;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_geometry(rsd_ctx* c, int is_period, int wall_method) {
  if (!c || (is_period != 0 && is_period != 1) || (wall_method != 0 && wall_method != 1)) return RSD_ERR_ARG;
  c->is_period = is_period;
  c->wall = wall_method;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_iterative(rsd_ctx* c, int precond, double tol, int max_iter, int inner_iters) {
  if (!c || precond < RSD_PRECOND_NONE || precond > RSD_PRECOND_MIXED_LEARNED || !(tol > 0.0) || !(tol < 1.0) ||
      max_iter < 1 || inner_iters < 1 || inner_iters > 31)
    return RSD_ERR_ARG;
  c->it_precond = precond;
  c->it_tol = tol;
  c->it_max = max_iter;
  c->it_inner = inner_iters;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_iterative_graphs(rsd_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return RSD_ERR_ARG;
  c->it_graphs = on != 0;
  return RSD_OK;
}

extern "C" int rsd_ctx_iterative_stats(rsd_ctx* c, int* iterations, double* relres, int* fell_back) {
  if (!c) return RSD_ERR_ARG;
  if (iterations) *iterations = c->last_it_iters;
  if (relres) *relres = c->last_it_relres;
  if (fell_back) *fell_back = c->last_it_fallback;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_resistance_route(rsd_ctx* c, int route) {
  if (!c || (route != RSD_RESIST_EXPLICIT && route != RSD_RESIST_SCHUR && route != RSD_RESIST_ITERATIVE))
    return RSD_ERR_ARG;
  c->resist = route;
  return RSD_OK;
}

extern "C" int rsd_ctx_substep_stats(rsd_ctx* c, int* cg_iterations, int* fallbacks) {
  if (!c) return RSD_ERR_ARG;
  if (cg_iterations) *cg_iterations = c->last_cg_iters;
  if (fallbacks) *fallbacks = c->last_fallbacks;
  return RSD_OK;
}

extern "C" int rsd_ctx_step_stats(rsd_ctx* c, int* cg_iterations, int* cg_fallbacks, int* explicit_redo) {
  if (!c) return RSD_ERR_ARG;
  if (cg_iterations) *cg_iterations = c->last_cg_iters;
  if (cg_fallbacks) *cg_fallbacks = c->last_fallbacks;
  if (explicit_redo) *explicit_redo = c->last_redo;
  return RSD_OK;
}

extern "C" int rsd_ctx_stage_times(rsd_ctx* c, double ms[6]) {
  if (!c || !ms) return RSD_ERR_ARG;
  for (int i = 0; i < 6; ++i) ms[i] = c->stage_ms[i];
  return RSD_OK;
}

// STEPPER_Stokesian (stepper.f90:33-156) on the periodic FTS path.
// What the host reads at the end of a step, gathered into its pinned copy by
// one thread: yeta_mu, INFO and the bounded-wait error words (a null word
// reads as 0); the one-launch CG's word is cleared once read.
// INFO: the grand mobility's (MobToResist_FTS's MATREV) when it failed, else
// that of the sub-steps' RPP factorisations
__global__ void step_end_kernel(const double* yeta, const int* g_info, const int* info, const int* solve_err,
                                const int* factor_err, unsigned* cg_err, const int* brown_info,
                                rsd_ctx::StepEnd* out) {
  if (threadIdx.x != 0) return;
  for (int q = 0; q < 5; ++q) out->yeta[q] = yeta[q];
  out->info = *g_info != 0x7fffffff ? *g_info : *info;
  out->brown_info = brown_info ? *brown_info : 0x7fffffff;
  out->wait_err[0] = solve_err ? *solve_err : 0;
  out->wait_err[1] = factor_err ? *factor_err : 0;
  out->cg_err = cg_err ? *cg_err : 0u;
  if (cg_err) *cg_err = 0u;
}


// ------------------------------------------------ the iterative route's system
struct rsd_gnn {
  GnnModel m;
};

static bool precond_learned(int p) { return p == RSD_PRECOND_LEARNED || p == RSD_PRECOND_MIXED_LEARNED; }
static bool precond_mixed(int p) { return p == RSD_PRECOND_MIXED || p == RSD_PRECOND_MIXED_LEARNED; }

// The FP32 copy of M: columns padded to 16 bytes for the vector loads
static int iter_ld32(int n11) { return (n11 + 3) & ~3; }

// Buffers of the iterative route (its Krylov space, the block-sparse
// lubrication, FP32 copies for the mixed preconditioners, the learned
// preconditioner's work), allocated on first use.
static int iter_buffers(rsd_ctx* c, int precond, int nz) {
  const int N = c->N, n11 = c->n11, cap = rsd_ctx::kBsrCap;
  RSD_TRY(iter_reserve(c->iw, n11, c->it_restart, N, cap));
  RSD_TRY(c->bsr_blk.alloc((size_t)N * cap * 121));
  RSD_TRY(c->bsr_self.alloc((size_t)N * 121));
  RSD_TRY(c->xit.alloc(2 * (size_t)n11));
  RSD_TRY(c->bit.alloc(2 * (size_t)n11));
  RSD_TRY(c->wit.alloc(2 * (size_t)n11));
  if (!c->bsr_ix) RSD_HIP(hipMalloc(&c->bsr_ix, ((size_t)N * (cap + 1) + 1) * sizeof(int)));
  const size_t n32 = (size_t)N * cap * 121 + (size_t)N * 121 + (size_t)iter_ld32(n11) * n11;
  if (precond_mixed(precond) && c->bsr32_cap < n32) {
    if (c->bsr32) RSD_HIP(hipFree(c->bsr32));
    c->bsr32 = nullptr;
    RSD_HIP(hipMalloc(&c->bsr32, n32 * sizeof(float)));
    c->bsr32_cap = n32;
  }
  // the pair distances feed the learned preconditioner's features
  RSD_TRY(gnn_work_reserve(c->gw, N, cap, std::max(nz, 1)));
  return RSD_OK;
}

static LubBsr<double> iter_bsr(rsd_ctx* c) {
  const int N = c->N, cap = rsd_ctx::kBsrCap;
  return LubBsr<double>{N, cap, c->bsr_ix, c->bsr_ix + N, c->bsr_blk.p, c->bsr_self.p};
}

// The lubrication's block-sparse rows at the current positions (lp: the
// step's lubrication parameters), block-Jacobi's inverses and, for a learned
// preconditioner, its correction (zero without a model); FP32 copies for the
// mixed ones.
static int iter_operator(rsd_ctx* c, const LubParams& lp, const rsdm::LubTables* tab, int precond, hipStream_t s) {
  const int N = c->N, cap = rsd_ctx::kBsrCap;
  const LubBsr<double> L = iter_bsr(c);
  LubParams lb = lp;
  lb.bsr_blk = c->bsr_blk.p;
  lb.bsr_self = c->bsr_self.p;
  lb.bsr_nnb = c->bsr_ix;
  lb.bsr_col = c->bsr_ix + N;
  lb.bsr_cap = cap;
  lb.bsr_overflow = c->bsr_ix + (size_t)N * (cap + 1);
  lb.bsr_dist = c->gw.dist;
  RSD_TRY(lub_add(lb, c->pos.p, c->radii.p, tab, c->flags + 1, s));
  RSD_TRY(block_jacobi_build(L, c->G.p, c->n11p, c->iw.dinv, s));
  if (precond_learned(precond)) {
    if (c->gnn) {
      RSD_TRY(gnn_prepare(c->gw, L, c->G.p, c->n11p, c->radii.p, c->iw.dinv, s));
      RSD_TRY(gnn_forward(c->gw, c->gnn->m.p, L, c->iw.corr, s));
    } else {
      RSD_TRY(launch_fill(c->iw.corr, (size_t)N * (cap + 1) * 121, 0.0, s));
    }
  }
  if (precond_mixed(precond)) {
    float* blk32 = c->bsr32;
    float* self32 = c->bsr32 + (size_t)N * cap * 121;
    RSD_TRY(to_f32(c->iw.dinv, c->iw.dinv32, (size_t)N * 121, s));
    RSD_TRY(to_f32(c->bsr_blk.p, blk32, (size_t)N * cap * 121, s));
    RSD_TRY(to_f32(c->bsr_self.p, self32, (size_t)N * 121, s));
  }
  return RSD_OK;
}

// GMRES's view of the system (the FP32 copies for the mixed preconditioners)
static GmresSetup iter_setup(rsd_ctx* c, int precond, const LubBsr<double>* L, LubBsr<float>* L32) {
  const int N = c->N, cap = rsd_ctx::kBsrCap, n11 = c->n11;
  const bool mixed = precond_mixed(precond);
  float* self32 = mixed ? c->bsr32 + (size_t)N * cap * 121 : nullptr;
  float* M32 = mixed ? self32 + (size_t)N * 121 : nullptr;
  *L32 = LubBsr<float>{N, cap, c->bsr_ix, c->bsr_ix + N, mixed ? c->bsr32 : nullptr, self32};
  GmresSetup gs{L, c->G.p, c->n11p, L32, M32, iter_ld32(n11), precond, c->it_inner, L};
  gs.graphs = c->it_graphs;
  return gs;
}

extern "C" int rsd_ctx_step(rsd_ctx* c, double dt, double t, double yeta_mu[5]) {
  if (!c || !c->has_state) {
    set_error("rsd_ctx_step: no state (call rsd_ctx_set_state)");
    return RSD_ERR_ARG;
  }
  c->iter_ready = false;
  rsdm::LubTables* tab = nullptr;
  RSD_TRY(device_tables(&tab));
  const int N = c->N, n6 = c->n6, n5 = c->n5, n11p = c->n11p, n6p = c->n6p;
  const rsd_step_config& cf = c->cfg;
  hipStream_t s = c->s;
  // INIT_u_bg (csd_post.f90:442-479)
  double ei_bg[6];
  memcpy(ei_bg, cf.ei_bg, sizeof ei_bg);
  ei_bg[4] = cf.gamma * 0.5;
  const double E[9] = {ei_bg[0], ei_bg[3], ei_bg[4], ei_bg[3], ei_bg[1], ei_bg[5], ei_bg[4], ei_bg[5], ei_bg[2]};
  double om[3] = {cf.omega_bg[0], cf.gamma * 0.5, cf.omega_bg[2]};
  // EI_transY2 (tensors.f90:890-903) with the Y2 basis
  const double s2 = 0.70710678118654752440, s6 = 0.40824829046386301637;
  StepEI ei;
  ei.v[0] = s2 * (E[1] + E[3]);
  ei.v[1] = -s2 * (E[5] + E[7]);
  ei.v[2] = -s6 * E[0] - s6 * E[4] + 2.0 * s6 * E[8];
  ei.v[3] = -s2 * (E[2] + E[6]);
  ei.v[4] = s2 * (E[0] - E[4]);
  StepKernelParams sp{};
  memcpy(sp.LB, cf.lb, sizeof sp.LB);
  sp.gamma = cf.gamma;
  sp.T = t;
  for (int I = 0; I < 3; ++I)
    for (int J = 0; J < 3; ++J) {
      // OmegaT(I,J) = sum_k eps(J,I,k) omega_bg(k)
      double o = 0.0;
      if (J == 1 && I == 0) o = -om[2];
      if (J == 2 && I == 0) o = om[1];
      if (J == 0 && I == 1) o = om[2];
      if (J == 2 && I == 1) o = -om[0];
      if (J == 0 && I == 2) o = -om[1];
      if (J == 1 && I == 2) o = om[0];
      sp.UOE[I + 3 * J] = o + E[I + 3 * J];
    }
  memcpy(sp.u_bg, cf.u_bg, sizeof sp.u_bg);
  memcpy(sp.omega_bg, om, sizeof sp.omega_bg);
  sp.N = N;
  sp.Nb = cf.nb;
  sp.wrap = c->is_period;

  RSD_HIP(hipEventRecord(c->ev[0], s));
  // The grand mobility's assembly goes first, so that the GPU runs it while
  // the host enqueues the step's small set-up work behind it (on this stream)
  // and beside it (on the side stream, which waits only for the clearing of
  // the flags: the overlap count the assembly adds to and the collision word)
  RSD_HIP(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), s));
  RSD_HIP(hipEventRecord(c->ev_fork, s));
  // SET_LATTICE_SKEW + CALC_LATTICE_INVERSE
  rsd_lattice lat;
  host_set_lattice_skew(t, cf.lb, cf.gamma, &lat);
  host_calc_lattice_inverse(&lat, cf.ews_err, N, c->h_radii.data());
  if (c->img_nr >= 0) {
    lat.nr = c->img_nr;
    lat.ni = c->img_ni;
  }
  const int side = 2 * lat.ni + 1;
  if (c->is_period) RSD_TRY(c->kw.alloc((size_t)side * side * side * rsdm::kKW));
  // MobToResist_FTS (matrices.f90:393-416): the resistance blocks of inv(G).
  // Explicit: W = inv(G), whose blocks the sub-steps copy.  Schur: with
  // G = [A B; B^T C] (A = APP, B = APQ, C = AQQ), the Cholesky factor of the
  // reordered [C B^T; B A] is [L_C 0; Y^T L_S] with Y = inv(L_C) B^T and
  // L_S L_S^T = S = A - B inv(C) B^T, so that
  //   W_PP = inv(S),  W_PQ EIN = -W_PP Y^T e  (e = inv(L_C) EIN),
  //   W_PQ^T x - W_QQ EIN = -inv(L_C)^T (Y (W_PP x - W_PQ EIN) + e)  (last sub-step).
  // The Schur route assembles the grand mobility straight into the reordered
  // layout (its Q fields first, each diagonal block padded to a multiple of
  // 64), which the factorisation then overwrites; a redo through the explicit
  // route assembles G in the reference's order again.
  const int n5p = c->n5p, npq = c->npq;
  double* Gq = c->W.p;                              // [C B^T; B A], then its factor
  const double* Yt = Gq + n5p;                      // Y^T: rows n5p.., columns 0..n5
  // PER_GRPERY_INV_FRI (periodic) or GRPERY_INV_FRI (free space) -> G
  // (padded with identity), or -> Gq; with Wall_method, Wall_GRPERY_INV_FRI's
  // TPP added to the TT block (stepper.f90:58-70)
  auto assemble = [&](bool reordered) -> int {
    EwaldParams ep = ewald_params(lat, N, n11p, cf.correction_method, cf.mu_f);
    ep.kmax2 = c->img_kmax2;
    double* dst = c->G.p;
    if (reordered) {
      ep.ld = npq;
      ep.offQ = 0;
      ep.offT = n5p;
      ep.offR = n5p + 3 * N;
      dst = Gq;
    }
    if (c->is_period) {
      RSD_TRY(ewald_assemble(ep, c->pos.p, c->radii.p, c->kw.p, dst, c->flags, s));
    } else {
      FreeParams fp{ep.inv_pimu, N, ep.ld, cf.correction_method, ep.offT, ep.offR, ep.offQ};
      RSD_TRY(free_assemble(fp, c->pos.p, c->radii.p, dst, c->flags, s));
    }
    if (c->wall) RSD_TRY(wall_tt(N, cf.mu_f, c->pos.p, c->radii.p, dst + ep.offT + (size_t)ep.offT * ep.ld, ep.ld, true, s));
    if (!reordered) return set_identity_padding(c->G.p, c->n11, n11p, n11p, s);
    return grand_padding(Gq, npq, n5, n5p, n6, n6p, s);
  };
  bool schur = c->resist == RSD_RESIST_SCHUR;
  const bool iterative = c->resist == RSD_RESIST_ITERATIVE;
  if (iterative && c->kbt > 0.0) {
    set_error("rsd_ctx_step: Brownian motion runs on the direct routes (its SijN_B needs the resistance RPQ)");
    return RSD_ERR_UNSUPPORTED;
  }
  RSD_TRY(assemble(schur));
  RSD_TRY(dense_info_reset(c->dw, s));
  RSD_TRY(dense_wait_reset(c->dw, s));
  if (!c->g_info) RSD_HIP(hipMalloc(&c->g_info, sizeof(int)));
  RSD_HIP(hipMemcpyAsync(c->g_info, c->dw.d_info, sizeof(int), hipMemcpyDeviceToDevice, s));
  RSD_TRY(launch_fill_ein(N, ei, c->ein.p, s));
  RSD_TRY(launch_fill_fe(N, cf.nb, cf.gravity, c->fe.p, s));
  // pos_collision_judge / ppiclf_collision_timestep (csd_post.f90:383-436), on
  // the side stream beside the grand mobility: the host learns NT_DEM from
  // the copy of flags[2] (ev[7], waited for before the sub-steps); nothing on
  // the main stream reads that word
  RSD_HIP(hipStreamWaitEvent(c->s2, c->ev_fork, 0));
  // the state a redo of the sub-steps starts from, saved on the side stream
  // too: nothing writes pos or upos before the first sub-step's Euler step,
  // which the host enqueues only after waiting for ev[7] below
  RSD_HIP(hipMemcpyAsync(c->pos_save.p, c->pos.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, c->s2));
  RSD_HIP(hipMemcpyAsync(c->upos_save.p, c->upos.p, 6 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, c->s2));
  RSD_TRY(launch_collision(N, cf.nb, c->pos.p, c->radii.p, c->flags + 2, c->s2));
  RSD_HIP(hipMemcpyAsync(c->h_flags + 2, c->flags + 2, sizeof(int), hipMemcpyDeviceToHost, c->s2));
  RSD_HIP(hipEventRecord(c->ev[7], c->s2));
  // averageU needs the mobility APP (matrices.f90:230-253): UU = APP F
  RSD_TRY(launch_fill_avgforce(N, 6.0 * kPi * cf.mu_f * cf.radii_test, c->F.p, s));
  if (schur) RSD_TRY(gemv(c->dw, Gq + n5p + (size_t)n5p * npq, npq, n6, n6, false, 1.0, c->F.p, 0.0, c->UU.p, s));
  else RSD_TRY(gemv(c->dw, c->G.p, n11p, n6, n6, false, 1.0, c->F.p, 0.0, c->UU.p, s));
  // BROWN (stepper.f90:72-86): A = CHOLESKY(APP(1:3N,1:3N)) (matrices.f90:126-157),
  // VB = sqrt(2 kBT) A z / DT, DMDX_CAL (stepper.f90:650-688): two more grand
  // mobilities at conf +- (deta/2) z, DMDX = (A1 - A2)(TT) z / deta;
  // u_brown = VB + kBT DMDX, which enters only SijN_B = RPQ^T UB (stepper.f90:407)
  if (c->kbt > 0.0) {
    const int n3 = 3 * N, n3p = round_up(n3, kTile), n11 = c->n11;
    const double deta = 0.001;
    RSD_TRY(c->bz.alloc(n3));
    RSD_TRY(c->ub.alloc(c->n6p));
    RSD_TRY(c->sb.alloc(n5));
    RSD_TRY(c->A3.alloc((size_t)n3p * n3p));
    RSD_TRY(c->Tb.alloc((size_t)n3 * n3));
    RSD_TRY(c->Gb.alloc((size_t)n11 * n11));
    RSD_TRY(c->bpos1.alloc(n3));
    RSD_TRY(c->bpos2.alloc(n3));
    if (!c->brown_info) RSD_HIP(hipMalloc(&c->brown_info, sizeof(int)));
    if (!c->z_next.empty()) {
      RSD_HIP(hipMemcpyAsync(c->bz.p, c->z_next.data(), n3 * sizeof(double), hipMemcpyHostToDevice, s));
      RSD_HIP(hipStreamSynchronize(s));  // z_next is pageable host memory reused by the next call
      c->z_next.clear();
    } else {
      RSD_TRY(launch_brown_normal(n3, c->brown_seed, c->brown_step, c->bz.p, s));
    }
    ++c->brown_step;
    const double* app = schur ? Gq + n5p + (size_t)n5p * npq : c->G.p;
    const int lda = schur ? npq : n11p;
    RSD_TRY(copy_block(app, lda, c->A3.p, n3p, n3, n3, s));
    RSD_TRY(set_identity_padding(c->A3.p, n3, n3p, n3p, s));
    RSD_TRY(potrf_lower(c->dw, c->A3.p, n3p, n3p, s));
    RSD_HIP(hipMemcpyAsync(c->brown_info, c->dw.d_info, sizeof(int), hipMemcpyDeviceToDevice, s));
    RSD_TRY(dense_info_reset(c->dw, s));
    RSD_TRY(zero_strict_upper(c->A3.p, n3p, n3p, s));
    RSD_TRY(launch_fill(c->ub.p, c->n6p, 0.0, s));
    RSD_TRY(gemv(c->dw, c->A3.p, n3p, n3, n3, false, sqrt(2.0 * c->kbt) / dt, c->bz.p, 0.0, c->ub.p, s));
    RSD_TRY(launch_brown_conf(N, c->pos.p, c->bz.p, deta, c->bpos1.p, c->bpos2.p, s));
    EwaldParams ep = ewald_params(lat, N, n11, cf.correction_method, cf.mu_f);
    ep.kmax2 = c->img_kmax2;
    for (int k = 0; k < 2; ++k) {
      const double* pk = k ? c->bpos2.p : c->bpos1.p;
      if (c->is_period) {
        RSD_TRY(ewald_assemble(ep, pk, c->radii.p, c->kw.p, c->Gb.p, c->flags + 3, s));
      } else {
        FreeParams fp{ep.inv_pimu, N, n11, cf.correction_method, 0, 3 * N, 6 * N};
        RSD_TRY(free_assemble(fp, pk, c->radii.p, c->Gb.p, c->flags + 3, s));
      }
      if (k == 0) RSD_TRY(copy_block(c->Gb.p, n11, c->Tb.p, n3, n3, n3, s));
      else RSD_TRY(matrix_axpy(-1.0, c->Gb.p, n11, c->Tb.p, n3, n3, n3, s));
    }
    RSD_TRY(gemv(c->dw, c->Tb.p, n3, n3, n3, false, c->kbt / deta, c->bz.p, 1.0, c->ub.p, s));
  }
  RSD_HIP(hipEventRecord(c->ev[1], s));
  // With the wall APP is not symmetric: averageU and BROWN's CHOLESKY (UPLO
  // 'L') above took it as it is; MobToResist_FTS's MATREV (UPLO 'U') inverts
  // its upper triangle mirrored, and so do the factorisations here (and the
  // iterative route's products)
  auto mirror_wall = [&](bool schur) -> int {
    if (!c->wall) return RSD_OK;
    const size_t ldg = schur ? npq : n11p, offt = schur ? n5p : 0;
    return mirror_upper(3 * N, (schur ? Gq : c->G.p) + offt + offt * ldg, ldg, s);
  };
  auto resist = [&](bool schur) -> int {
    RSD_TRY(mirror_wall(schur));
    if (!schur) return spd_inverse(c->dw, c->G.p, c->W.p, n11p, n11p, s);
    // the factor, and for systems of up to kFusedInverseMaxTiles block
    // columns X = inv(L_S) in the same launch, formed by tiles that run
    // behind the diagonal (into RPP's buffer, free until the first sub-step
    // rebuilds RPP). There the launch is bound by its diagonal chain and the
    // X tiles fill its idle capacity; in larger systems the tile work bounds
    // it, and the separate triangular inverse after it is faster.
    const bool fused_inverse = npq / 64 <= kFusedInverseMaxTiles;
    if (fused_inverse) {
      RSD_TRY(potrf_lower_x(c->dw, Gq, npq, npq, n5p / 64, c->RPP.p, n6p, s));
    } else {
      RSD_TRY(potrf_lower(c->dw, Gq, npq, npq, s));
    }
    // e = inv(L_C) EIN (rows past n5 zero) and Y^T e need only the factor:
    // they run on the side stream while the main stream forms W_PP (which
    // touches neither L_C, Y, dinvC nor the triangular-solve and GEMV scratch)
    RSD_HIP(hipEventRecord(c->ev_fork, s));
    RSD_HIP(hipStreamWaitEvent(c->s2, c->ev_fork, 0));
    // L_C's diagonal inverses outlive this step's other factorisations (the
    // main stream's next write of w.dinv, the first sub-step's, follows ev_join)
    RSD_HIP(hipMemcpyAsync(c->dinvC.p, c->dw.dinv, (size_t)n5p * 64 * sizeof(double), hipMemcpyDeviceToDevice, c->s2));
    RSD_TRY(launch_fill(c->ec.p, n5p, 0.0, c->s2));
    RSD_HIP(hipMemcpyAsync(c->ec.p, c->ein.p, n5 * sizeof(double), hipMemcpyDeviceToDevice, c->s2));
    RSD_TRY(trsv_lower(c->dw, Gq, n5p, npq, c->dinvC.p, c->ec.p, false, c->s2));
    RSD_TRY(gemv(c->dw, Yt, npq, n6, n5, false, 1.0, c->ec.p, 0.0, c->v6.p, c->s2));
    RSD_HIP(hipEventRecord(c->ev_join, c->s2));
    // W_PP = inv(L_S)^T inv(L_S)
    if (fused_inverse) {
      RSD_TRY(lauum_lower(c->dw, c->RPP.p, n6p, c->WPP.p, n6p, n6p, s));
    } else {
      // L_S, then inv(L_S) in place; neither it nor its diagonal inverses
      // overlap L_C, Y^T or L_C's diagonal inverses, which the side stream reads
      double* LS = Gq + n5p + (size_t)n5p * npq;
      RSD_TRY(trtri_lower(c->dw, LS, n6p, npq, s, c->dw.dinv + (size_t)n5p * 64));
      RSD_TRY(lauum_lower(c->dw, LS, npq, c->WPP.p, n6p, n6p, s));
    }
    RSD_TRY(symmetrize_from_lower(c->WPP.p, n6p, n6p, s));
    // W_PQ EIN = -W_PP Y^T e
    RSD_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
    RSD_TRY(gemv(c->dw, c->WPP.p, n6p, n6, n6, false, -1.0, c->v6.p, 0.0, c->wpq_ein.p, s));
    return RSD_OK;
  };
  if (!iterative) {
    RSD_TRY(resist(schur));
    RSD_HIP(hipMemcpyAsync(c->g_info, c->dw.d_info, sizeof(int), hipMemcpyDeviceToDevice, s));
  } else {
    RSD_TRY(mirror_wall(false));
  }
  RSD_HIP(hipEventRecord(c->ev[2], s));
  RSD_HIP(hipEventSynchronize(c->ev[7]));
  int nt = 1;
  if (c->h_flags[2]) nt = (dt < cf.dt_dem0) ? 5 : std::min((int)std::floor(dt / cf.dt_dem0), 100);
  if (nt < 1) nt = 1;
  c->last_nt = nt;
  sp.dt = dt / nt;
  LubParams lp{};
  memcpy(lp.LR, lat.lr, sizeof lp.LR);
  memcpy(lp.LI, lat.li, sizeof lp.LI);
  lp.scale = 6.0 * (kPi * cf.mu_f);
  lp.N = N;
  lp.Nb = cf.nb;
  lp.is_period = c->is_period;
  lp.PP = c->RPP.p;
  lp.ldPP = n6p;
  lp.PQ = c->RPQ.p;
  lp.ldPQ = n6;
  lp.QQ = c->RQQ.p;
  lp.ldQQ = n5;
  const bool two_rhs = cf.gravity != 0.0;
  double* x0 = c->upar.p;          // U_par, then (one right-hand side) also inv(RPP) RPQ EIN
  double* x1 = c->upar.p + n6p;    // inv(RPP) RPQ EIN when Fe is nonzero
  double* b0 = c->bvec.p;          // their right-hand sides, for the CG sub-steps
  double* b1 = c->bvec.p + n6p;
  const int nrhs = two_rhs ? 2 : 1;
  const bool use_cg = nt > 1 && c->solver != RSD_SUBSTEP_DIRECT && (nt - 1) * nrhs <= kPcgMaxSlots;
  c->last_cg_iters = 0;
  c->last_fallbacks = 0;
  // RPQ EIN (+ Fe) into y0 (and RPQ EIN into y1): explicit, RPQ holds W_PQ + lubrication;
  // Schur, W_PQ EIN comes from the step's vector and the lubrication's part
  // from lub_add (RPQ is not formed)
  auto rhs = [&](bool sch, double* y0, double* y1) -> int {
    // y1 (or y0 with one right-hand side) = RPQ EIN; y0 = Fe + RPQ EIN
    double* ye = two_rhs ? y1 : y0;
    if (sch) RSD_TRY(vector_sum(c->wpq_ein.p, c->lpe.p, ye, n6, s));
    else RSD_TRY(gemv(c->dw, c->RPQ.p, n6, n6, n5, false, 1.0, c->ein.p, 0.0, ye, s));
    if (two_rhs) RSD_TRY(vector_sum(c->fe.p, ye, y0, n6, s));
    return RSD_OK;
  };
  // The NT_DEM sub-steps (stepper.f90:95-143).  With cg, sub-steps after the
  // first are solved by CG (enqueued without waiting); *nslots gets the
  // number of CG solves, whose status slots are read once all are enqueued.
  auto substeps = [&](bool cg, bool sch, int* nslots) -> int {
    bool have_precond = false, copy_pending = false, precond_pending = false;
    // Schur: RPQ and RQQ hold only the lubrication, which enters the step
    // through RPQ EIN, RQQ EIN and RPQ^T x: lub_add forms those from the pairs
    lp.ein = sch ? c->ein.p : nullptr;
    lp.pq_ein = sch ? c->lpe.p : nullptr;
    lp.qq_ein = sch ? c->lqe.p : nullptr;
    int slot = 0;
    for (int sub = 0; sub < nt; ++sub) {
      // RPP is rebuilt below: the side stream's copy of the first factor goes first
      if (copy_pending) {
        RSD_HIP(hipStreamWaitEvent(s, c->ev_copied, 0));
        copy_pending = false;
      }
      // arrangResist: R = inv(M) + lubrication (hydro_lub.f90 lubmxcalc with the current positions)
      if (sch) {
        RSD_TRY(copy_block(c->WPP.p, n6p, c->RPP.p, n6p, n6p, n6p, s));
      } else {
        RSD_TRY(copy_block(c->W.p, n11p, c->RPP.p, n6p, n6, n6, s));
        RSD_TRY(set_identity_padding(c->RPP.p, n6, n6p, n6p, s));
        RSD_TRY(copy_block(c->W.p + (size_t)n6 * n11p, n11p, c->RPQ.p, n6, n6, n5, s));
        RSD_TRY(copy_block(c->W.p + n6 + (size_t)n6 * n11p, n11p, c->RQQ.p, n5, n5, n5, s));
      }
      RSD_TRY(lub_add(lp, c->pos.p, c->radii.p, tab, c->flags + 1, s));
      if (sub == 0) RSD_HIP(hipEventRecord(c->ev[3], s));
      // STEPPER_NOGmres (stepper.f90:387-403) needs inv(RPP) only applied to
      // vectors: U_par = inv(RPP) (Fe + RPQ EIN) and
      // SijN_hyd = RPQ^T inv(RPP) RPQ EIN - RQQ EIN, so RPP is factored and
      // solved, or (later sub-steps with cg) solved by CG preconditioned with
      // the first sub-step's inverse, from the previous two sub-steps' solutions
      // extrapolated.
      if (cg && have_precond) {
        RSD_TRY(rhs(sch, b0, b1));
        RSD_TRY(pcg_initial_guess(x0, c->xhist.p, nrhs * n6p, slot == 0, s));
        if (precond_pending) {
          RSD_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
          precond_pending = false;
        }
        for (int j = 0; j < nrhs; ++j, ++slot)
          RSD_TRY(pcg_solve(c->pw, c->RPP.p, n6p, n6p, j ? b1 : b0, j ? x1 : x0, c->cg_tol, c->cg_max_iter, slot,
                            c->solver == RSD_SUBSTEP_PCG_ONE_LAUNCH, c->cg_pred[slot] + 2, s));
      } else if (cg) {
        // the first sub-step: its factor solves it (two triangular solves).
        // On the side stream meanwhile, inv(RPP) is formed in FP32 from an
        // FP32 copy of the factor, to precondition the later sub-steps (that
        // stream's scratch, w.tmp and the GEMM descriptors, is not the
        // triangular solves' or the GEMVs' on this stream)
        RSD_TRY(potrf_lower(c->dw, c->RPP.p, n6p, n6p, s));
        RSD_HIP(hipEventRecord(c->ev_fork, s));
        RSD_HIP(hipStreamWaitEvent(c->s2, c->ev_fork, 0));
        RSD_TRY(pcg_precond_from_factor(c->dw, c->pw, c->RPP.p, n6p, n6p, c->s2, c->ev_copied));
        RSD_HIP(hipEventRecord(c->ev_join, c->s2));
        RSD_TRY(rhs(sch, x0, x1));
        RSD_TRY(potrs_lower(c->dw, c->RPP.p, n6p, n6p, x0, n6p, nrhs, s));
        have_precond = copy_pending = precond_pending = true;
      } else {
        RSD_TRY(potrf_lower(c->dw, c->RPP.p, n6p, n6p, s));
        RSD_TRY(rhs(sch, x0, x1));
        RSD_TRY(potrs_lower(c->dw, c->RPP.p, n6p, n6p, x0, n6p, nrhs, s));
      }
      if (sub == 0) RSD_HIP(hipEventRecord(c->ev[4], s));
      // SijN_hyd of the last sub-step is the one the step keeps (each sub-step
      // overwrites it, stepper.f90:133-141)
      if (sub == nt - 1) {
        const double* xs = two_rhs ? x1 : x0;
        if (sch) {
          // RPQ^T x from the pairs again (the positions are this sub-step's), - RQQ EIN
          LubParams lt = lp;
          lt.ein = nullptr;
          lt.pq_ein = lt.qq_ein = nullptr;
          lt.xt = xs;
          lt.pqt_x = c->shyd.p;
          RSD_TRY(lub_add(lt, c->pos.p, c->radii.p, tab, nullptr, s));
          RSD_TRY(matrix_axpy(-1.0, c->lqe.p, n5, c->shyd.p, n5, n5, 1, s));
        } else {
          RSD_TRY(gemv(c->dw, c->RPQ.p, n6, n6, n5, true, 1.0, xs, 0.0, c->shyd.p, s));
          RSD_TRY(gemv(c->dw, c->RQQ.p, n5, n5, n5, false, -1.0, c->ein.p, 1.0, c->shyd.p, s));
        }
        if (sch) {
          // + W_PQ^T x - W_QQ EIN = -inv(L_C)^T (Y (W_PP x - W_PQ EIN) + e)
          RSD_HIP(hipMemcpyAsync(c->v6.p, c->wpq_ein.p, n6 * sizeof(double), hipMemcpyDeviceToDevice, s));
          RSD_TRY(gemv(c->dw, c->WPP.p, n6p, n6, n6, false, 1.0, xs, -1.0, c->v6.p, s));
          RSD_HIP(hipMemcpyAsync(c->v5.p, c->ec.p, n5p * sizeof(double), hipMemcpyDeviceToDevice, s));
          RSD_TRY(gemv(c->dw, Yt, npq, n6, n5, true, 1.0, c->v6.p, 1.0, c->v5.p, s));
          RSD_TRY(trsv_lower(c->dw, Gq, n5p, npq, c->dinvC.p, c->v5.p, true, s));
          RSD_TRY(matrix_axpy(-1.0, c->v5.p, n5, c->shyd.p, n5, n5, 1, s));
        }
        // SijN_B = RPQ^T UB (stepper.f90:407): Schur, the lubrication's part
        // from the pairs and W_PQ^T UB = -inv(L_C)^T Y W_PP UB
        if (c->kbt > 0.0) {
          if (sch) {
            LubParams lb = lp;
            lb.ein = nullptr;
            lb.pq_ein = lb.qq_ein = nullptr;
            lb.xt = c->ub.p;
            lb.pqt_x = c->sb.p;
            RSD_TRY(lub_add(lb, c->pos.p, c->radii.p, tab, nullptr, s));
            RSD_TRY(gemv(c->dw, c->WPP.p, n6p, n6, n6, false, 1.0, c->ub.p, 0.0, c->v6.p, s));
            RSD_TRY(launch_fill(c->v5.p, n5p, 0.0, s));
            RSD_TRY(gemv(c->dw, Yt, npq, n6, n5, true, 1.0, c->v6.p, 0.0, c->v5.p, s));
            RSD_TRY(trsv_lower(c->dw, Gq, n5p, npq, c->dinvC.p, c->v5.p, true, s));
            RSD_TRY(matrix_axpy(-1.0, c->v5.p, n5, c->sb.p, n5, n5, 1, s));
          } else {
            RSD_TRY(gemv(c->dw, c->RPQ.p, n6, n6, n5, true, 1.0, c->ub.p, 0.0, c->sb.p, s));
          }
        }
      }
      // STEPPER_conf: Euler step and Lees-Edwards wrap with the pre-step t
      RSD_TRY(launch_euler(sp, c->pos.p, x0, c->upos.p, s));
    }
    *nslots = slot;
    return RSD_OK;
  };
  // The iterative route: each sub-step solves K x = [Fe; -EIN] by GMRES
  // (rsd_iter.hip), U = (M x)_P and, at the last sub-step, SijN_hyd =
  // (x + R_lub M x)_Q (with gravity, a second solve with Fe = 0 gives SijN_hyd,
  // which excludes the force's part, as the reference's mu EIN does).
  // A solve that does not converge redoes the step on the direct Schur route.
  auto substeps_iter = [&]() -> int {
    const int n11 = c->n11, cap = rsd_ctx::kBsrCap, pc = c->it_precond;
    RSD_TRY(iter_buffers(c, pc, 1));
    const LubBsr<double> L = iter_bsr(c);
    LubBsr<float> L32;
    const GmresSetup gs = iter_setup(c, pc, &L, &L32);
    if (precond_mixed(pc)) RSD_TRY(matrix_to_f32(c->G.p, n11p, const_cast<float*>(gs.M32), gs.ld32, n11, n11, s));
    RSD_TRY(launch_fill_int(c->bsr_ix + (size_t)N * (cap + 1), 1, 0, s));
    RSD_TRY(launch_fill(c->xit.p, 2 * (size_t)n11, 0.0, s));
    c->last_it_iters = 0;
    c->last_it_relres = 0.0;
    for (int sub = 0; sub < nt; ++sub) {
      RSD_TRY(iter_operator(c, lp, tab, pc, s));
      if (sub == 0) RSD_HIP(hipEventRecord(c->ev[3], s));
      const int nsolve = (two_rhs && sub == nt - 1) ? 2 : 1;
      for (int k = 0; k < nsolve; ++k) {
        double* b = c->bit.p + (size_t)k * n11;
        double* x = c->xit.p + (size_t)k * n11;
        RSD_TRY(launch_iter_rhs(N, k == 0 ? c->fe.p : nullptr, c->ein.p, b, s));
        GmresResult res;
        RSD_TRY(gmres_solve(c->iw, gs, b, x, c->it_tol, c->it_max, &res, s));
        c->last_it_iters += res.iterations;
        c->last_it_relres = std::max(c->last_it_relres, res.relres);
        if (!res.converged) return 1;
      }
      // U = (M x)_P
      RSD_TRY(dense_gemv<double>(c->G.p, n11p, n11, c->xit.p, c->wit.p, c->iw.part, s));
      RSD_HIP(hipMemcpyAsync(x0, c->wit.p, n6 * sizeof(double), hipMemcpyDeviceToDevice, s));
      if (sub == 0) RSD_HIP(hipEventRecord(c->ev[4], s));
      if (sub == nt - 1) {
        const double* xs = c->xit.p + (size_t)(nsolve - 1) * n11;
        RSD_TRY(dense_gemv<double>(c->G.p, n11p, n11, xs, c->wit.p, c->iw.part, s));
        RSD_TRY(bsr_apply<double>(L, c->wit.p, c->wit.p + n11, false, s));
        RSD_TRY(vector_sum(xs + n6, c->wit.p + n11 + n6, c->shyd.p, n5, s));
      }
      RSD_TRY(launch_euler(sp, c->pos.p, x0, c->upos.p, s));
    }
    return RSD_OK;
  };
  int nslots = 0;
  c->last_it_fallback = 0;
  if (iterative) {
    const int rc = substeps_iter();
    if (rc < 0) return rc;
    if (rc == 1) {
      // GMRES did not converge: the direct Schur route redoes the step from its start
      RSD_HIP(hipMemcpyAsync(c->pos.p, c->pos_save.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
      RSD_HIP(hipMemcpyAsync(c->upos.p, c->upos_save.p, 6 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
      RSD_HIP(hipStreamSynchronize(s));
      c->resist = RSD_RESIST_SCHUR;
      const int rc2 = rsd_ctx_step(c, dt, t, yeta_mu);
      c->resist = RSD_RESIST_ITERATIVE;
      c->last_it_fallback = 1;
      return rc2;
    }
    int ovf = 0;
    RSD_HIP(hipMemcpyAsync(&ovf, c->bsr_ix + (size_t)N * (rsd_ctx::kBsrCap + 1), sizeof(int), hipMemcpyDeviceToHost, s));
    RSD_HIP(hipStreamSynchronize(s));
    if (ovf) {
      set_error("rsd_ctx_step: a particle has more lubrication neighbours than the block-sparse rows hold");
      return RSD_ERR_CAPACITY;
    }
  } else {
    RSD_TRY(substeps(use_cg, schur, &nslots));
  }
  // One wait at the end: yeta_mu, INFO, the CG statuses and the bounded-wait
  // error words.  A CG solve that did not converge redoes the sub-steps with
  // factorisations; a failed factorisation on the Schur route redoes the
  // step's dense part through the explicit inverse, whose INFO is the
  // reference's (the order of G's first non-positive leading minor).
  rsd_ctx::StepEnd& he = *c->h_end;
  auto finish = [&]() -> int {
    RSD_HIP(hipEventRecord(c->ev[5], s));
    RSD_TRY(launch_yeta(N, c->shyd.p, c->kbt > 0.0 ? c->sb.p : nullptr, c->ein.p, c->UU.p, lat.lv, cf.mu_f, lat.phi,
                        c->yeta.p, s));
    RSD_HIP(hipEventRecord(c->ev[6], s));
    // yeta_mu, INFO and the error words into the pinned StepEnd in one small
    // launch (the CG statuses are in pinned memory already)
    hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(64), 0, s, (const double*)c->yeta.p, (const int*)c->g_info,
                       (const int*)c->dw.d_info,
                       dense_solve_error_word(c->dw), dense_factor_error_word(c->dw), pcg_error_word(c->pw),
                       (const int*)(c->kbt > 0.0 ? c->brown_info : nullptr), c->d_end);
    RSD_HIP(hipGetLastError());
    RSD_HIP(hipStreamSynchronize(s));
    if (he.cg_err) {
      set_error("rsd_ctx_step: a conjugate-gradient workgroup ran out of its bounded barrier wait");
      return RSD_ERR_HIP;
    }
    if (he.wait_err[0] || he.wait_err[1]) {
      std::string m = std::string("rsd_ctx_step: a workgroup of the dense ") +
                      (he.wait_err[1] ? "factorisation" : "triangular solves") + " ran out of its bounded wait";
      int rec[4];
      if (he.wait_err[1] && dense_factor_wait_record(c->dw, rec) == RSD_OK && rec[0])
        m += " (ticket " + std::to_string(rec[1]) + " waited on flag " + std::to_string(rec[2]) + ", which held " +
             std::to_string(rec[3]) + ")";
      set_error(m);
      return RSD_ERR_HIP;
    }
    if (he.brown_info != 0x7fffffff) {
      set_error("rsd_ctx_step: Brownian CHOLESKY of APP(1:3N,1:3N): the leading minor of order " +
                std::to_string(he.brown_info) + " is not positive definite (the reference stops, matrices.f90:140-147)");
      return RSD_ERR_NOT_SPD;
    }
    return RSD_OK;
  };
  RSD_TRY(finish());
  for (int i = 0; i < nslots; ++i) {
    const int* st = c->pw.h_state + 2 * i;
    c->last_cg_iters += st[1];
    if (st[0] != kPcgConverged) ++c->last_fallbacks;
    else c->cg_pred[i] = st[1];
  }
  const bool refactor = schur && he.info != 0x7fffffff;
  c->last_redo = refactor ? 1 : 0;
  if (c->last_fallbacks || refactor) {
    RSD_HIP(hipMemcpyAsync(c->pos.p, c->pos_save.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
    RSD_HIP(hipMemcpyAsync(c->upos.p, c->upos_save.p, 6 * (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
    RSD_HIP(hipMemsetAsync(c->flags + 1, 0, sizeof(int), s));
    if (refactor) {
      RSD_TRY(dense_info_reset(c->dw, s));
      schur = false;
      RSD_HIP(hipMemsetAsync(c->flags, 0, sizeof(int), s));  // the overlap count, counted again
      RSD_TRY(assemble(false));
      RSD_TRY(resist(false));
      RSD_HIP(hipMemcpyAsync(c->g_info, c->dw.d_info, sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    RSD_TRY(substeps(false, schur, &nslots));
    RSD_TRY(finish());
  }
  double y[5];
  memcpy(y, he.yeta, sizeof y);
  const int info = he.info;
  float ms;
  const int pairs[5][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 6}};
  for (int k = 0; k < 5; ++k) {
    RSD_HIP(hipEventElapsedTime(&ms, c->ev[pairs[k][0]], c->ev[pairs[k][1]]));
    c->stage_ms[k] = ms;
  }
  RSD_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[6]));
  c->stage_ms[5] = ms;
  if (yeta_mu) memcpy(yeta_mu, y, sizeof y);
  return (info == 0x7fffffff) ? RSD_OK : info;
}

// ------------------------------------------------ the learned preconditioner
extern "C" int rsd_gnn_create(rsd_gnn** out, unsigned long long seed) {
  if (!out) return RSD_ERR_ARG;
  std::unique_ptr<rsd_gnn> g(new rsd_gnn);
  RSD_TRY(gnn_model_init(g->m, seed));
  *out = g.release();
  return RSD_OK;
}

extern "C" int rsd_gnn_destroy(rsd_gnn* g) {
  if (!g) return RSD_ERR_ARG;
  gnn_model_release(g->m);
  delete g;
  return RSD_OK;
}

extern "C" int rsd_gnn_get_params(rsd_gnn* g, float* params) {
  if (!g || !params) return RSD_ERR_ARG;
  RSD_HIP(hipMemcpy(params, g->m.p, kGnnParams * sizeof(float), hipMemcpyDeviceToHost));
  return RSD_OK;
}

extern "C" int rsd_gnn_set_params(rsd_gnn* g, const float* params) {
  if (!g || !params) return RSD_ERR_ARG;
  RSD_HIP(hipMemcpy(g->m.p, params, kGnnParams * sizeof(float), hipMemcpyHostToDevice));
  RSD_HIP(hipMemset(g->m.m, 0, kGnnParams * sizeof(float)));
  RSD_HIP(hipMemset(g->m.v, 0, kGnnParams * sizeof(float)));
  g->m.t = 0;
  return RSD_OK;
}

extern "C" int rsd_ctx_set_gnn(rsd_ctx* c, rsd_gnn* g) {
  if (!c) return RSD_ERR_ARG;
  c->gnn = g;
  return RSD_OK;
}

static_assert(RSD_GNN_PARAMS == kGnnParams, "the header's parameter count");

extern "C" int rsd_ctx_prepare_iterative(rsd_ctx* c, double t) {
  if (!c || !c->has_state) {
    set_error("rsd_ctx_prepare_iterative: no state (call rsd_ctx_set_state)");
    return RSD_ERR_ARG;
  }
  rsdm::LubTables* tab = nullptr;
  RSD_TRY(device_tables(&tab));
  const int N = c->N, n11p = c->n11p, cap = rsd_ctx::kBsrCap;
  const rsd_step_config& cf = c->cfg;
  hipStream_t s = c->s;
  c->iter_ready = false;
  // EIN as the step forms it (INIT_u_bg, EI_transY2 with the Y2 basis)
  double ei_bg[6];
  memcpy(ei_bg, cf.ei_bg, sizeof ei_bg);
  ei_bg[4] = cf.gamma * 0.5;
  const double E[9] = {ei_bg[0], ei_bg[3], ei_bg[4], ei_bg[3], ei_bg[1], ei_bg[5], ei_bg[4], ei_bg[5], ei_bg[2]};
  const double s2 = 0.70710678118654752440, s6 = 0.40824829046386301637;
  StepEI ei;
  ei.v[0] = s2 * (E[1] + E[3]);
  ei.v[1] = -s2 * (E[5] + E[7]);
  ei.v[2] = -s6 * E[0] - s6 * E[4] + 2.0 * s6 * E[8];
  ei.v[3] = -s2 * (E[2] + E[6]);
  ei.v[4] = s2 * (E[0] - E[4]);
  RSD_TRY(launch_fill_ein(N, ei, c->ein.p, s));
  RSD_TRY(launch_fill_int(c->flags, 4, 0, s));
  // the grand mobility (as the step assembles it for the explicit route)
  rsd_lattice lat;
  host_set_lattice_skew(t, cf.lb, cf.gamma, &lat);
  host_calc_lattice_inverse(&lat, cf.ews_err, N, c->h_radii.data());
  if (c->img_nr >= 0) {
    lat.nr = c->img_nr;
    lat.ni = c->img_ni;
  }
  const int side = 2 * lat.ni + 1;
  EwaldParams ep = ewald_params(lat, N, n11p, cf.correction_method, cf.mu_f);
  ep.kmax2 = c->img_kmax2;
  if (c->is_period) {
    RSD_TRY(c->kw.alloc((size_t)side * side * side * rsdm::kKW));
    RSD_TRY(ewald_assemble(ep, c->pos.p, c->radii.p, c->kw.p, c->G.p, c->flags, s));
  } else {
    FreeParams fp{ep.inv_pimu, N, ep.ld, cf.correction_method, ep.offT, ep.offR, ep.offQ};
    RSD_TRY(free_assemble(fp, c->pos.p, c->radii.p, c->G.p, c->flags, s));
  }
  if (c->wall) {
    RSD_TRY(wall_tt(N, cf.mu_f, c->pos.p, c->radii.p, c->G.p, n11p, true, s));
    RSD_TRY(mirror_upper(3 * N, c->G.p, n11p, s));
  }
  RSD_TRY(set_identity_padding(c->G.p, c->n11, n11p, n11p, s));
  // the block-sparse lubrication, block-Jacobi and the pair pieces
  RSD_TRY(iter_buffers(c, RSD_PRECOND_MIXED_LEARNED, 8));
  LubParams lp{};
  memcpy(lp.LR, lat.lr, sizeof lp.LR);
  memcpy(lp.LI, lat.li, sizeof lp.LI);
  lp.scale = 6.0 * (kPi * cf.mu_f);
  lp.N = N;
  lp.Nb = cf.nb;
  lp.is_period = c->is_period;
  RSD_TRY(launch_fill_int(c->bsr_ix + (size_t)N * (cap + 1), 1, 0, s));
  RSD_TRY(iter_operator(c, lp, tab, RSD_PRECOND_BLOCK_JACOBI, s));
  const LubBsr<double> L = iter_bsr(c);
  RSD_TRY(gnn_prepare(c->gw, L, c->G.p, n11p, c->radii.p, c->iw.dinv, s));
  int ovf = 0;
  RSD_HIP(hipMemcpyAsync(&ovf, c->bsr_ix + (size_t)N * (cap + 1), sizeof(int), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipStreamSynchronize(s));
  if (ovf) {
    set_error("rsd_ctx_prepare_iterative: a particle has more lubrication neighbours than the block-sparse rows hold");
    return RSD_ERR_CAPACITY;
  }
  c->iter_ready = true;
  return RSD_OK;
}

static int need_prepared(rsd_ctx* c, const char* who) {
  if (!c) return RSD_ERR_ARG;
  if (!c->iter_ready) {
    set_error(std::string(who) + ": no prepared system (call rsd_ctx_prepare_iterative)");
    return RSD_ERR_ARG;
  }
  return RSD_OK;
}

extern "C" int rsd_ctx_iterative_system(rsd_ctx* c, double* M, double* R, int* nnb, int* col, double* dist,
                                        int* cap_out) {
  RSD_TRY(need_prepared(c, "rsd_ctx_iterative_system"));
  const int N = c->N, n11 = c->n11, cap = rsd_ctx::kBsrCap;
  hipStream_t s = c->s;
  if (cap_out) *cap_out = cap;
  if (M) RSD_TRY(d2h_block(c->G.p, c->n11p, M, n11, n11, n11, s));
  std::vector<int> hn(N), hc((size_t)N * cap);
  RSD_HIP(hipMemcpyAsync(hn.data(), c->bsr_ix, N * sizeof(int), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipMemcpyAsync(hc.data(), c->bsr_ix + N, (size_t)N * cap * sizeof(int), hipMemcpyDeviceToHost, s));
  RSD_HIP(hipStreamSynchronize(s));
  if (nnb) memcpy(nnb, hn.data(), N * sizeof(int));
  if (col) memcpy(col, hc.data(), (size_t)N * cap * sizeof(int));
  if (dist) RSD_HIP(hipMemcpy(dist, c->gw.dist, (size_t)N * cap * sizeof(double), hipMemcpyDeviceToHost));
  if (R) {
    std::vector<double> blk((size_t)N * cap * 121), self((size_t)N * 121);
    RSD_HIP(hipMemcpy(blk.data(), c->bsr_blk.p, blk.size() * sizeof(double), hipMemcpyDeviceToHost));
    RSD_HIP(hipMemcpy(self.data(), c->bsr_self.p, self.size() * sizeof(double), hipMemcpyDeviceToHost));
    auto dof = [N](int g, int k) { return k < 3 ? 3 * g + k : (k < 6 ? 3 * N + 3 * g + k - 3 : 6 * N + 5 * g + k - 6); };
    std::fill(R, R + (size_t)n11 * n11, 0.0);
    for (int g = 0; g < N; ++g) {
      for (int r = 0; r < 11; ++r)
        for (int k = 0; k < 11; ++k) R[(size_t)dof(g, r) + (size_t)dof(g, k) * n11] = self[(size_t)g * 121 + r + 11 * k];
      for (int q = 0; q < hn[g]; ++q) {
        const int d = hc[(size_t)g * cap + q];
        const double* b = &blk[((size_t)g * cap + q) * 121];
        for (int r = 0; r < 11; ++r)
          for (int k = 0; k < 11; ++k) R[(size_t)dof(g, r) + (size_t)dof(d, k) * n11] = b[r + 11 * k];
      }
    }
  }
  return RSD_OK;
}

extern "C" int rsd_ctx_gnn_loss(rsd_ctx* c, rsd_gnn* g, int nz, unsigned long long seed, double* loss, float* grad,
                                double* z) {
  RSD_TRY(need_prepared(c, "rsd_ctx_gnn_loss"));
  if (!g || !loss || nz < 1 || nz > c->gw.nz) return RSD_ERR_ARG;
  hipStream_t s = c->s;
  const LubBsr<double> L = iter_bsr(c);
  RSD_TRY(gnn_forward(c->gw, g->m.p, L, c->iw.corr, s));
  RSD_TRY(gnn_loss_grad(c->gw, g->m.p, L, c->G.p, c->n11p, c->iw.dinv, c->iw.corr, nz, seed, 0, grad != nullptr,
                        false, c->iw.part, loss, s));
  if (grad) RSD_HIP(hipMemcpy(grad, c->gw.grad, kGnnParams * sizeof(float), hipMemcpyDeviceToHost));
  if (z) RSD_HIP(hipMemcpy(z, c->gw.zv, (size_t)nz * c->n11 * sizeof(double), hipMemcpyDeviceToHost));
  return RSD_OK;
}

extern "C" int rsd_ctx_gnn_train(rsd_ctx* c, rsd_gnn* g, int iters, int nz, float lr, unsigned long long seed,
                                 double* losses) {
  RSD_TRY(need_prepared(c, "rsd_ctx_gnn_train"));
  if (!g || iters < 0 || nz < 1 || nz > c->gw.nz || !(lr > 0.0f)) return RSD_ERR_ARG;
  hipStream_t s = c->s;
  const LubBsr<double> L = iter_bsr(c);
  for (int it = 0; it < iters; ++it) {
    double loss = 0.0;
    RSD_TRY(gnn_forward(c->gw, g->m.p, L, c->iw.corr, s));
    RSD_TRY(gnn_loss_grad(c->gw, g->m.p, L, c->G.p, c->n11p, c->iw.dinv, c->iw.corr, nz, seed,
                          (unsigned long long)it + 1, true, true, c->iw.part, &loss, s));
    RSD_TRY(gnn_adam(g->m, c->gw.grad, lr, s));
    if (losses) losses[it] = loss;
  }
  RSD_HIP(hipStreamSynchronize(s));
  return RSD_OK;
}

extern "C" int rsd_ctx_iterative_solve(rsd_ctx* c, int precond, double tol, int max_iter, int* iterations,
                                       double* relres, double* ms) {
  RSD_TRY(need_prepared(c, "rsd_ctx_iterative_solve"));
  if (precond < RSD_PRECOND_NONE || precond > RSD_PRECOND_MIXED_LEARNED || !(tol > 0.0) || max_iter < 1)
    return RSD_ERR_ARG;
  const int N = c->N, n11 = c->n11, cap = rsd_ctx::kBsrCap;
  hipStream_t s = c->s;
  RSD_HIP(hipEventRecord(c->ev[0], s));
  const LubBsr<double> L = iter_bsr(c);
  if (precond_learned(precond)) {
    if (c->gnn) RSD_TRY(gnn_forward(c->gw, c->gnn->m.p, L, c->iw.corr, s));
    else RSD_TRY(launch_fill(c->iw.corr, (size_t)N * (cap + 1) * 121, 0.0, s));
  }
  LubBsr<float> L32;
  const GmresSetup gs = iter_setup(c, precond, &L, &L32);
  if (precond_mixed(precond)) {
    RSD_TRY(matrix_to_f32(c->G.p, c->n11p, const_cast<float*>(gs.M32), gs.ld32, n11, n11, s));
    RSD_TRY(to_f32(c->iw.dinv, c->iw.dinv32, (size_t)N * 121, s));
    RSD_TRY(to_f32(c->bsr_blk.p, c->bsr32, (size_t)N * cap * 121, s));
    RSD_TRY(to_f32(c->bsr_self.p, c->bsr32 + (size_t)N * cap * 121, (size_t)N * 121, s));
  }
  RSD_TRY(launch_iter_rhs(N, nullptr, c->ein.p, c->bit.p, s));
  RSD_TRY(launch_fill(c->xit.p, n11, 0.0, s));
  GmresResult res;
  RSD_TRY(gmres_solve(c->iw, gs, c->bit.p, c->xit.p, tol, max_iter, &res, s));
  RSD_HIP(hipEventRecord(c->ev[1], s));
  RSD_HIP(hipEventSynchronize(c->ev[1]));
  float t_ms = 0.0f;
  RSD_HIP(hipEventElapsedTime(&t_ms, c->ev[0], c->ev[1]));
  if (iterations) *iterations = res.iterations;
  if (relres) *relres = res.relres;
  if (ms) *ms = t_ms;
  return RSD_OK;
}

// ------------------------------------------------------------------ batches
// Many independent suspensions stepped together: each context keeps its own
// streams, and `width` host threads take the contexts (largest first) from a
// shared counter, so that the steps of small suspensions, which fill a
// fraction of the GPU each, overlap on it.  The arithmetic of every step is
// the context's own: a batch gives the same numbers as stepping each alone.
struct rsd_batch {
  std::vector<rsd_ctx*> ctx;
  std::vector<int> order;  // contexts by decreasing N
  int width = 1;
};

extern "C" int rsd_batch_create(rsd_batch** out, int count, rsd_ctx* const* ctxs) {
  if (!out || count < 1 || !ctxs) return RSD_ERR_ARG;
  std::unique_ptr<rsd_batch> b(new rsd_batch);
  for (int i = 0; i < count; ++i) {
    if (!ctxs[i]) return RSD_ERR_ARG;
    for (rsd_ctx* c : b->ctx)
      if (c == ctxs[i]) {
        set_error("rsd_batch_create: a context appears twice");
        return RSD_ERR_ARG;
      }
    b->ctx.push_back(ctxs[i]);
    b->order.push_back(i);
  }
  std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return b->ctx[x]->N > b->ctx[y]->N; });
  b->width = std::min(count, 8);
  *out = b.release();
  return RSD_OK;
}

extern "C" int rsd_batch_destroy(rsd_batch* b) {
  if (!b) return RSD_ERR_ARG;
  delete b;
  return RSD_OK;
}

extern "C" int rsd_batch_set_width(rsd_batch* b, int width) {
  if (!b || width < 1 || width > (int)b->ctx.size() || width > 256) return RSD_ERR_ARG;
  b->width = width;
  return RSD_OK;
}

extern "C" int rsd_batch_get_width(rsd_batch* b, int* width) {
  if (!b || !width) return RSD_ERR_ARG;
  *width = b->width;
  return RSD_OK;
}

extern "C" int rsd_batch_step(rsd_batch* b, double dt, const double* t, double* yeta_mu, int* info) {
  if (!b) return RSD_ERR_ARG;
  if (prof_on()) {
    set_error("rsd_batch_step: kernel timing is on (rsd_profile_enable); its records are not shared between threads");
    return RSD_ERR_ARG;
  }
  const int n = (int)b->ctx.size(), W = std::min(b->width, n);
  std::atomic<int> next{0};
  std::vector<int> rcs(n, RSD_OK);
  std::vector<std::string> msgs(n);
  std::vector<double> ybuf(5 * (size_t)n);
  auto work = [&]() {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= n) break;
      const int i = b->order[k];
      rsd_ctx* c = b->ctx[i];
      if (hipSetDevice(c->dev) != hipSuccess) {
        rcs[i] = RSD_ERR_HIP;
        msgs[i] = "rsd_batch_step: hipSetDevice failed";
        continue;
      }
      rcs[i] = rsd_ctx_step(c, dt, t ? t[i] : 0.0, &ybuf[5 * (size_t)i]);
      if (rcs[i] < 0) msgs[i] = rsd_last_error();
    }
  };
  std::vector<std::thread> th;
  th.reserve(W - 1);
  for (int w = 1; w < W; ++w) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
  for (int i = 0; i < n; ++i)
    if (rcs[i] < 0) {
      set_error("rsd_batch_step: suspension " + std::to_string(i) + ": " + msgs[i]);
      return rcs[i];
    }
  if (yeta_mu) memcpy(yeta_mu, ybuf.data(), ybuf.size() * sizeof(double));
  if (info)
    for (int i = 0; i < n; ++i) info[i] = rcs[i];
  return RSD_OK;
}

// The width (host threads) that steps the batch fastest, among `ncand`
// candidates: every context's state is saved, one batch step warms up
// (first-step allocations), then for each width the states are restored and
// one batch step is timed; the states are restored at the end and the best
// width kept.  ms[ncand] (may be NULL) receives each candidate's time.
extern "C" int rsd_batch_autotune(rsd_batch* b, double dt, const double* t, const int* cand, int ncand, int* best,
                                  double* ms) {
  if (!b || !cand || ncand < 1) return RSD_ERR_ARG;
  const int n = (int)b->ctx.size();
  std::vector<std::vector<double>> conf(n), upos(n), radii(n);
  for (int i = 0; i < n; ++i) {
    conf[i].resize(3 * (size_t)b->ctx[i]->N);
    upos[i].resize(6 * (size_t)b->ctx[i]->N);
    radii[i] = b->ctx[i]->h_radii;
    RSD_TRY(rsd_ctx_get_state(b->ctx[i], conf[i].data(), upos[i].data()));
  }
  auto restore = [&]() -> int {
    for (int i = 0; i < n; ++i) RSD_TRY(rsd_ctx_set_state(b->ctx[i], conf[i].data(), radii[i].data(), upos[i].data()));
    return RSD_OK;
  };
  const int keep = b->width;
  RSD_TRY(rsd_batch_step(b, dt, t, nullptr, nullptr));
  double best_ms = 1e300;
  int best_w = keep;
  for (int k = 0; k < ncand; ++k) {
    const int w = std::max(1, std::min(cand[k], n));
    RSD_TRY(restore());
    b->width = w;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = rsd_batch_step(b, dt, t, nullptr, nullptr);
    const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc < 0) {
      b->width = keep;
      return rc;
    }
    if (ms) ms[k] = el;
    if (el < best_ms) {
      best_ms = el;
      best_w = w;
    }
  }
  RSD_TRY(restore());
  b->width = best_w;
  if (best) *best = best_w;
  return RSD_OK;
}

extern "C" int rsd_stepper_stokesian(int n, double* conf, const double* radii, double dt, double t,
                                     const rsd_step_config* cfg, double* yeta_mu, double* u_pos) {
  if (n < 1 || !conf || !radii || !cfg) return RSD_ERR_ARG;
  rsd_ctx* c = nullptr;
  RSD_TRY(rsd_ctx_create(&c, n, cfg));
  std::unique_ptr<rsd_ctx, int (*)(rsd_ctx*)> guard(c, rsd_ctx_destroy);
  RSD_TRY(rsd_ctx_set_state(c, conf, radii, u_pos));
  int rc = rsd_ctx_step(c, dt, t, yeta_mu);
  if (rc < 0) return rc;
  RSD_TRY(rsd_ctx_get_state(c, conf, u_pos));
  return rc;
}
