// rsd_run -- drop-in for the reference's main program (main.f90:3-254) on the
// periodic FTS path, stepping on the MI355X library.
//
// Run it where the reference would run: it reads control_file.dat and
// initial_config.dat from the current directory (INIT_READ_CONTROL,
// INITIAL_SIZE, INITIAL_CONF), the lubrication tables from lubdat/ (or the
// directory given as the first argument), nondimensionalises as
// dimentionless_calc/_translate, and then runs main.f90's time loop:
//   every K_WRITE steps a VMD_WRITE frame into output.VTF before the step and
//   a yeta_mu.txt row after it (main.f90:178-201), T += DT until T > END_TIME,
//   and a stop with an error when U_pos(2) is NaN (main.f90:223-226).
// A hot start (initial_config.dat with hot_start true and rows
// T,x,y,z,a,Ux,Uy,Uz, the rows output.VTF holds) resumes a run.
//
// Settings outside the periodic FTS path (no Ewald images, GMRES, DEM
// stepper, walls, boundaries, Brownian motion, oscillating shear,
// conglomerates, filaments, lattice-generated configurations) stop the run
// with exit code 2.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rsd.h"

static int fail(const char* what, int rc) {
  fprintf(stderr, "rsd_run: %s failed (%d): %s\n", what, rc, rsd_last_error());
  return 1;
}

int main(int argc, char** argv) {
  const char* lubdat = argc > 1 ? argv[1] : "lubdat";
  rsd_control ctl;
  int rc = rsd_read_control("control_file.dat", &ctl);
  if (rc) return fail("reading control_file.dat", rc);
  struct {
    bool bad;
    const char* what;
  } unsupported[] = {
      {!ctl.is_period || !ctl.simple_period, "IsPeriod and simplePeriod must be true (periodic Lees-Edwards box)"},
      {!ctl.fts_method, "FTS_method must be true"},
      {!ctl.uselub, "uselub must be true"},
      {ctl.use_dem || ctl.use_dem_stepper, "the DEM packing and DEM stepper are outside this path"},
      {ctl.use_gmres != 0, "useGMRESmethod: the iterative route is selected in the library, not here"},
      {ctl.oscillation_shear || ctl.fix || ctl.boundary || ctl.wall_method || ctl.brown,
       "oscillating shear, fixed particles, boundaries, walls and Brownian motion are outside this path"},
      {ctl.is_case_lattice != 0, "lattice-generated configurations (IsCaseLattice != 0) are outside this path"},
      {ctl.correction_method < 0 || ctl.correction_method > 2, "correction_method must be 0, 1 or 2"},
      {ctl.k_write < 1, "K_WRITE must be positive"},
  };
  for (const auto& u : unsupported)
    if (u.bad) {
      fprintf(stderr, "rsd_run: %s\n", u.what);
      return 2;
    }
  rsd_units un;
  rc = rsd_nondimensionalise(&ctl, &un);
  if (rc) return fail("nondimensionalising", rc);
  int n = 0, hot = 0;
  double lb[3], T = 0.0;
  rc = rsd_read_initial_config("initial_config.dat", 0, &n, lb, nullptr, nullptr, nullptr, nullptr, &hot);
  if (rc) return fail("reading initial_config.dat", rc);
  std::vector<double> conf(3 * (size_t)n), radii(n), u_pos(6 * (size_t)n, 0.0);
  T = un.t0;  // INITIAL_CONF starts from T0 and takes each row's T
  rc = rsd_read_initial_config("initial_config.dat", n, &n, lb, &T, conf.data(), radii.data(), u_pos.data(), &hot);
  if (rc) return fail("reading initial_config.dat", rc);

  if ((rc = rsd_device_init(0))) return fail("rsd_device_init", rc);
  if ((rc = rsd_lub_tables_load(lubdat))) return fail("loading the lubrication tables", rc);
  rsd_step_config cfg;
  rsd_step_config_default(&cfg, lb[0]);
  memcpy(cfg.lb, lb, sizeof cfg.lb);
  cfg.gamma = un.gamma;
  cfg.ews_err = ctl.linerr;
  cfg.mu_f = un.mu_f;
  cfg.correction_method = ctl.correction_method;
  cfg.nb = 0;
  cfg.gravity = un.gravity;
  cfg.dt_dem0 = un.dt_dem0;
  cfg.radii_test = ctl.radii_test;
  memcpy(cfg.u_bg, ctl.u_bg, sizeof cfg.u_bg);
  memcpy(cfg.omega_bg, ctl.omega_bg, sizeof cfg.omega_bg);
  memcpy(cfg.ei_bg, ctl.ei_bg, sizeof cfg.ei_bg);
  rsd_ctx* ctx = nullptr;
  if ((rc = rsd_ctx_create(&ctx, n, &cfg))) return fail("rsd_ctx_create", rc);
  if ((rc = rsd_ctx_set_state(ctx, conf.data(), radii.data(), u_pos.data()))) return fail("rsd_ctx_set_state", rc);

  const double dt = un.base_dt, gamma_angle = 0.0;
  FILE* t1 = fopen("output.VTF", "w");
  FILE* t2 = fopen("yeta_mu.txt", "w");
  if (!t1 || !t2) {
    fprintf(stderr, "rsd_run: cannot create output.VTF / yeta_mu.txt\n");
    return 1;
  }
  fclose(t1);
  fclose(t2);
  double yeta[5] = {0, 0, 0, 0, 0};
  for (long k = 0;; ++k) {
    const bool write = k % ctl.k_write == 0;
    if (write) {
      if ((rc = rsd_ctx_get_state(ctx, conf.data(), u_pos.data()))) return fail("rsd_ctx_get_state", rc);
      // VMD_WRITE takes U_pos's translations: 3 entries per particle (initial_conf.f90:255-256)
      if ((rc = rsd_write_vtf_frame("output.VTF", 1, n, 0, conf.data(), T / un.lambda_time, radii.data(),
                                    u_pos.data())))
        return fail("writing output.VTF", rc);
    }
    rc = rsd_ctx_step(ctx, dt, T, yeta);
    if (rc < 0) return fail("rsd_ctx_step", rc);
    if (rc > 0) fprintf(stderr, "rsd_run: step %ld: the grand mobility is not positive definite (INFO %d)\n", k, rc);
    if (write &&
        (rc = rsd_write_yeta_mu("yeta_mu.txt", 1, T / un.lambda_time, yeta, lb[0], un.frequency, un.gamma,
                                gamma_angle)))
      return fail("writing yeta_mu.txt", rc);
    T += dt;
    if (T > un.end_time) break;
    if ((rc = rsd_ctx_get_state(ctx, nullptr, u_pos.data()))) return fail("rsd_ctx_get_state", rc);
    if (std::isnan(u_pos[1])) {
      fprintf(stderr, "rsd_run: U_pos has error %g %g %g\n", u_pos[0], u_pos[1], u_pos[2]);
      rsd_ctx_destroy(ctx);
      return 1;
    }
  }
  rsd_ctx_destroy(ctx);
  return 0;
}
