// rsd_io.cpp -- the reference's on-disk formats and nondimensionalisation
// (host only; part of librsd_amd.so).
//
//   control_file.dat    INIT_READ_CONTROL   initial_conf.f90:109-189
//   initial_config.dat  INITIAL_SIZE        initial_conf.f90:356-438
//                       INITIAL_CONF        initial_conf.f90:442-476 (IsCaseLattice = 0, hot start)
//   output.VTF          VMD_WRITE           initial_conf.f90:250-262
//   yeta_mu.txt         main.f90:197-201 (= yetamu_WRITE, initial_conf.f90:278-292)
//   units               dimentionless_calc / _translate  mymodules.f90:808-881,
//                       frequency and GAMMA as main.f90:148-155
//
// Both input files are read the way Fortran list-directed READ(unit,*) reads
// them: each READ takes as many values as it names from the next record(s),
// then drops the rest of its last record (so trailing "!" remarks vanish);
// reals accept the D exponent letter; logicals are .true./T/.false./F.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rsd.h"

namespace rsd {
void set_error(const std::string& msg);  // rsd_capi.hip: the text rsd_last_error() returns
constexpr double kPi = 3.141592653589793238462643383279502884;

namespace {

constexpr double kPai2 = 9.8696044010893586188344909998762;  // mymodules.f90:48

// One Fortran list-directed input stream over a text file.
class ListReader {
 public:
  explicit ListReader(FILE* f) : f_(f) {}

  // READ(u,*) with no items: skip one record.
  bool skip_record() {
    std::string line;
    return next_line(&line);
  }
  // READ(u,*) v1..vk: k values, across records if need be; the rest of the
  // last record is dropped.
  bool read(std::vector<std::string>* vals, int k) {
    vals->clear();
    std::string line;
    while ((int)vals->size() < k) {
      if (!next_line(&line)) return false;
      size_t i = 0;
      while ((int)vals->size() < k && i < line.size()) {
        while (i < line.size() && (isspace((unsigned char)line[i]) || line[i] == ',')) ++i;
        if (i >= line.size()) break;
        if (line[i] == '/') return (int)vals->size() == k;  // slash ends the READ
        size_t j = i;
        while (j < line.size() && !isspace((unsigned char)line[j]) && line[j] != ',' && line[j] != '/') ++j;
        vals->push_back(line.substr(i, j - i));
        i = j;
      }
    }
    return true;
  }

 private:
  bool next_line(std::string* line) {
    line->clear();
    int c;
    bool any = false;
    while ((c = fgetc(f_)) != EOF) {
      any = true;
      if (c == '\n') break;
      if (c != '\r') line->push_back((char)c);
    }
    return any;
  }
  FILE* f_;
};

bool to_real(const std::string& s, double* v) {
  std::string t = s;
  for (char& ch : t)
    if (ch == 'd' || ch == 'D') ch = 'e';
  char* end = nullptr;
  errno = 0;
  *v = strtod(t.c_str(), &end);
  return end && *end == '\0' && errno == 0 && !t.empty();
}

bool to_int(const std::string& s, int* v) {
  char* end = nullptr;
  errno = 0;
  long x = strtol(s.c_str(), &end, 10);
  if (!end || *end != '\0' || errno || s.empty()) return false;
  *v = (int)x;
  return true;
}

// Fortran logical input: optional '.', then T or F (case-insensitive).
bool to_logical(const std::string& s, int* v) {
  size_t i = 0;
  if (i < s.size() && s[i] == '.') ++i;
  if (i >= s.size()) return false;
  const char c = (char)toupper((unsigned char)s[i]);
  if (c != 'T' && c != 'F') return false;
  *v = c == 'T';
  return true;
}

struct Item {
  char kind;  // 'i' integer, 'l' logical, 'r' real
  void* p;
  int count;
  const char* name;
};

int read_items(ListReader& rd, const std::vector<Item>& items, const std::string& path, int* line_no) {
  std::vector<std::string> vals;
  for (const Item& it : items) {
    ++*line_no;
    if (!rd.read(&vals, it.count)) {
      set_error(path + ": ends before " + it.name + " (READ " + std::to_string(*line_no) + ")");
      return RSD_ERR_ARG;
    }
    for (int k = 0; k < it.count; ++k) {
      bool ok = it.kind == 'i'   ? to_int(vals[k], static_cast<int*>(it.p) + k)
                : it.kind == 'l' ? to_logical(vals[k], static_cast<int*>(it.p) + k)
                                 : to_real(vals[k], static_cast<double*>(it.p) + k);
      if (!ok) {
        set_error(path + ": cannot read " + it.name + " from '" + vals[k] + "' (READ " + std::to_string(*line_no) +
                  ")");
        return RSD_ERR_ARG;
      }
    }
  }
  return RSD_OK;
}

// Fortran Ew.dEe output (E24.15E3 here): 0.d1..dd, signed exponent of e digits.
void fortran_e(std::string* out, double v, int w, int d, int e) {
  char buf[64];
  if (v == 0.0) {
    snprintf(buf, sizeof buf, "%s0.%0*dE+%0*d", std::signbit(v) ? "-" : "", d, 0, e, 0);
  } else {
    char m[64];
    snprintf(m, sizeof m, "%.*E", d - 1, fabs(v));  // d1.d2..dd E+-X: value = 0.d1..dd x 10^(X+1)
    const char* ep = strchr(m, 'E');
    const int x = atoi(ep + 1) + 1;
    std::string digits;
    for (const char* q = m; q < ep; ++q)
      if (isdigit((unsigned char)*q)) digits.push_back(*q);
    snprintf(buf, sizeof buf, "%s0.%sE%c%0*d", v < 0 ? "-" : "", digits.c_str(), x < 0 ? '-' : '+', e, abs(x));
  }
  const int len = (int)strlen(buf);
  if (len > w) {
    out->append(w, '*');
    return;
  }
  out->append(w - len, ' ');
  out->append(buf);
}

// Fortran ESw.d output (ES24.15): d1.d2..E+-XX, or +-XXX without the E
// letter once the exponent needs three digits.
void fortran_es(std::string* out, double v, int w, int d) {
  char m[64], buf[80];
  snprintf(m, sizeof m, "%.*E", d, v);
  char* ep = strchr(m, 'E');
  const int x = atoi(ep + 1);
  *ep = '\0';
  if (abs(x) <= 99)
    snprintf(buf, sizeof buf, "%sE%c%02d", m, x < 0 ? '-' : '+', abs(x));
  else
    snprintf(buf, sizeof buf, "%s%c%03d", m, x < 0 ? '-' : '+', abs(x));
  const int len = (int)strlen(buf);
  if (len > w) {
    out->append(w, '*');
    return;
  }
  out->append(w - len, ' ');
  out->append(buf);
}

}  // namespace
}  // namespace rsd

using namespace rsd;

extern "C" int rsd_read_control(const char* path, rsd_control* c) {
  if (!path || !c) return RSD_ERR_ARG;
  FILE* f = fopen(path, "r");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return RSD_ERR_ARG;
  }
  memset(c, 0, sizeof *c);
  ListReader rd(f);
  const std::vector<Item> items = {
      {'i', &c->correction_method, 1, "correction_method"}, {'l', &c->fts_method, 1, "FTS_method"},
      {'l', &c->filament_solve_implicit, 1, "filament_solve_implicit"}, {'l', &c->uselub, 1, "uselub"},
      {'l', &c->use_dem, 1, "useDEM"}, {'l', &c->usecollision, 1, "usecollision"},
      {'l', &c->use_dem_stepper, 1, "useDEMstepper"}, {'l', &c->use_gmres, 1, "useGMRESmethod"},
      {'l', &c->xshear, 1, "Xshear"}, {'l', &c->is_period, 1, "IsPeriod"},
      {'l', &c->simple_period, 1, "simplePeriod"}, {'l', &c->oscillation_shear, 1, "Oscillation_shear"},
      {'l', &c->fix, 1, "fix"}, {'i', &c->is_case_lattice, 1, "IsCaseLattice"},
      {'i', &c->beginstep, 1, "beginstep"}, {'i', &c->interval, 1, "interval"},
      {'i', &c->num_threads, 1, "NUM_THREADS"}, {'i', &c->timestep, 1, "timestep"},
      {'r', &c->radii_test, 1, "radii_test"}, {'r', &c->radii_true, 1, "radii_true"},
      {'r', &c->g_gravity, 1, "g_gravity"}, {'r', c->u_bg, 3, "u_bg"}, {'r', c->omega_bg, 3, "omega_bg"},
      {'r', c->ei_bg, 6, "EI_bg"}, {'r', &c->mu_f, 1, "mu_f"}, {'r', &c->pho_f, 1, "pho_f"},
      {'r', &c->re_number, 1, "Re_number"}, {'r', &c->gamma_alpha, 1, "Gamma_alpha"},
      {'r', &c->youngs_module, 1, "Youngs_module"}, {'r', &c->poisson_ratio, 1, "Poisson_ratio"},
      {'r', &c->ksp, 1, "ksp"}, {'r', &c->erest, 1, "erest"}, {'r', &c->pho_par, 1, "pho_par"},
      {'r', &c->linerr, 1, "LINERR"}, {'i', &c->collision_condition, 1, "collision_condition"},
      {'l', &c->mu_calculate, 1, "muCalculate"}, {'r', &c->t0, 1, "T0"}, {'r', &c->end_time, 1, "END_TIME"},
      {'r', &c->base_dt, 1, "BASE_DT"}, {'i', &c->k_write, 1, "K_WRITE"}, {'r', &c->lj_eps, 1, "LJ_EPS"},
      {'r', &c->lj_sigma, 1, "LJ_SIGMA"}, {'r', &c->lj_cut, 1, "LJ_CUT"}, {'l', &c->boundary, 1, "Boundary"},
      {'l', &c->wall_method, 1, "wall_method"}, {'l', &c->use_dlvo, 1, "useDLVO"},
      {'r', &c->debye_length, 1, "DebyeLength"}, {'r', &c->z0, 1, "z0"}, {'r', &c->lambda_l, 1, "Lambda_l"},
      {'r', &c->permittivity, 1, "Permittivity"}, {'r', &c->potential, 1, "Potential"}, {'r', &c->ah, 1, "AH"},
      {'r', &c->surfaceenergy, 1, "Surfaceenergy"}, {'l', &c->brown, 1, "BROWN"}};
  int line_no = 0;
  int rc = read_items(rd, items, path, &line_no);
  if (!rc && c->brown) rc = read_items(rd, {{'r', &c->kbt, 1, "kBT"}}, path, &line_no);
  fclose(f);
  return rc;
}

extern "C" int rsd_read_initial_config(const char* path, int cap, int* n, double lb[3], double* t, double* conf,
                                       double* radii, double* u_pos, int* hot_start) {
  if (!path || !n) return RSD_ERR_ARG;
  FILE* f = fopen(path, "r");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return RSD_ERR_ARG;
  }
  ListReader rd(f);
  int nxyz[3], np = 0, krb[2], frb[2], hs = 0, line_no = 0;
  double ang[6], L[3];
  int rc = read_items(rd, {{'i', nxyz, 3, "NX,NY,NZ"}, {'i', &np, 1, "Np"}, {'i', krb, 2, "K_rb,Nswimer"}}, path,
                      &line_no);
  if (!rc && krb[0] != 0) {
    set_error(std::string(path) + ": rigid conglomerates (K_rb != 0) are outside the periodic FTS path");
    rc = RSD_ERR_UNSUPPORTED;
  }
  if (!rc) rd.skip_record();  // the KK_rbmconn line (empty when K_rb = 0)
  if (!rc) rc = read_items(rd, {{'i', frb, 2, "F_rb,Nfilament"}}, path, &line_no);
  if (!rc && frb[0] != 0) {
    set_error(std::string(path) + ": filaments (F_rb != 0) are outside the periodic FTS path");
    rc = RSD_ERR_UNSUPPORTED;
  }
  if (!rc) rd.skip_record();  // the Filament_num line
  if (!rc)
    rc = read_items(rd, {{'r', ang, 6, "alphaX,betaY,gamaZ,DX,DY,DZ"}, {'r', L, 3, "LB"}, {'l', &hs, 1, "hot_start"}},
                    path, &line_no);
  if (!rc && np < 1) {
    set_error(std::string(path) + ": Np must be positive");
    rc = RSD_ERR_ARG;
  }
  if (!rc) {
    *n = np;
    if (lb) memcpy(lb, L, sizeof L);
    if (hot_start) *hot_start = hs;
  }
  if (!rc && conf) {
    if (np > cap || !radii) {
      set_error(std::string(path) + ": more particles than the arrays hold");
      rc = RSD_ERR_CAPACITY;
    }
    double T = 0.0;
    for (int i = 0; i < np && !rc; ++i) {
      double row[8] = {0};
      rc = read_items(rd, {{'r', row, hs ? 8 : 5, "particle row"}}, path, &line_no);
      if (rc) break;
      T = row[0];
      for (int k = 0; k < 3; ++k) conf[3 * i + k] = row[1 + k];
      radii[i] = row[4];
      if (hs && u_pos)
        for (int k = 0; k < 3; ++k) u_pos[3 * i + k] = row[5 + k];
    }
    if (!rc && t) *t = T;  // the last row's T, as INITIAL_CONF leaves it
  }
  fclose(f);
  return rc;
}

extern "C" int rsd_nondimensionalise(const rsd_control* c, rsd_units* u) {
  if (!c || !u) return RSD_ERR_ARG;
  if (!(c->mu_f > 0.0) || !(c->radii_true > 0.0) || !(c->radii_test > 0.0)) {
    set_error("rsd_nondimensionalise: mu_f, radii_test and radii_true must be positive");
    return RSD_ERR_ARG;
  }
  memset(u, 0, sizeof *u);
  // dimentionless_calc (mymodules.f90:808-835)
  const double l_renold = 1.0, l_kb = 1.0, l_velocity = 1.0;
  const double l_visc = 1.0 / kPi / c->mu_f;
  const double l_len = c->radii_test / c->radii_true;
  const double l_pho = l_renold * l_visc / (l_velocity * l_len);
  const double l_mass = l_pho * l_len * l_len * l_len;
  const double l_time = l_len / l_velocity;
  const double l_acc = l_len / (l_time * l_time);
  const double l_force = l_mass * l_acc;
  const double l_temp = l_len * l_visc * l_visc / (l_pho * l_kb);
  const double l_press = l_force / (l_len * l_len);
  const double l_ksp = l_press * l_len;
  // dimentionless_translate (mymodules.f90:837-881), the quantities the step reads
  const double ksp = c->ksp * l_ksp;
  const double pho_par = c->pho_par * l_pho;
  const double mass_par = pho_par * 4 * kPi / 3.0 * 1.0;
  u->lambda_time = l_time;
  u->mu_f = c->mu_f * l_visc;
  u->gravity = mass_par * c->g_gravity * l_acc;
  u->end_time = c->end_time * l_time;
  u->base_dt = c->base_dt * l_time;
  u->t0 = c->t0 * l_time;
  u->dt_dem0 = sqrt(mass_par / ksp * (log(c->erest) * log(c->erest) + kPai2)) / 10.0;
  u->kbt = c->kbt * l_kb * l_temp;
  // main.f90:148-155 (steady shear)
  u->frequency = 1.0 / l_time;
  u->gamma = c->gamma_alpha * u->frequency;
  return RSD_OK;
}

static int write_text(const char* path, int append, const std::string& text) {
  FILE* f = fopen(path, append ? "a" : "w");
  if (!f) {
    set_error(std::string("cannot open ") + path + " for writing");
    return RSD_ERR_ARG;
  }
  const size_t w = fwrite(text.data(), 1, text.size(), f);
  const int bad = fclose(f) != 0 || w != text.size();
  if (bad) {
    set_error(std::string("short write to ") + path);
    return RSD_ERR_ARG;
  }
  return RSD_OK;
}

extern "C" int rsd_write_vtf_frame(const char* path, int append, int n, int nb, const double* conf, double t,
                                   const double* radii, const double* u_pos) {
  if (!path || n < 1 || nb < 0 || nb > n || !conf || !radii || !u_pos) return RSD_ERR_ARG;
  std::string text;
  for (int i = 0; i < n - nb; ++i) {
    const double row[8] = {t, conf[3 * i], conf[3 * i + 1], conf[3 * i + 2], radii[i],
                           u_pos[3 * i], u_pos[3 * i + 1], u_pos[3 * i + 2]};
    for (double v : row) fortran_e(&text, v, 24, 15, 3);
    text.push_back('\n');
  }
  return write_text(path, append, text);
}

extern "C" int rsd_write_yeta_mu(const char* path, int append, double t, const double yeta_mu[5], double lb1,
                                 double frequency, double gamma, double gamma_angle) {
  if (!path || !yeta_mu) return RSD_ERR_ARG;
  std::string text;
  const double row[10] = {t, yeta_mu[0], yeta_mu[1], yeta_mu[2], yeta_mu[3], yeta_mu[4], lb1, frequency, gamma,
                          gamma_angle};
  for (double v : row) fortran_es(&text, v, 24, 15);
  text.push_back('\n');
  return write_text(path, append, text);
}
