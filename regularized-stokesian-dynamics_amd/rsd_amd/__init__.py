"""rsd_amd -- Python binding of the MI355X regularized Stokesian dynamics library.

A thin ctypes layer over the C ABI declared in ``include/rsd.h``
(``librsd_amd.so`` built from ``csrc/`` by ``__graft_entry__.build()`` or
``make -C regularized-stokesian-dynamics_amd/csrc``).  Function names and
argument meanings follow the reference Fortran routines they replace
(``per_GRPY_YZD.f90``, ``hydro_lub.f90``, ``matrices.f90``, ``stepper.f90``).

There is no CPU fallback: importing works without a GPU, but every compute
call needs the HIP library and a device, and raises otherwise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "librsd_amd.so")
LUBDAT_DIR = os.path.join(PKG_DIR, "data", "lubdat")
GNN_PARAMS_PATH = os.path.join(PKG_DIR, "data", "gnn_params.npy")  # the trained learned preconditioner
PAI = np.pi

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

RSD_OK = 0
ERRORS = {-1: "RSD_ERR_ARG", -2: "RSD_ERR_HIP", -3: "RSD_ERR_NOMEM", -4: "RSD_ERR_TABLES",
          -5: "RSD_ERR_CAPACITY", -6: "RSD_ERR_NODEVICE", -7: "RSD_ERR_UNSUPPORTED", -8: "RSD_ERR_NOT_SPD"}


class RSDError(RuntimeError):
    pass


class Lattice(C.Structure):
    """``rsd_lattice``: the reference's LATTICE_BASE state (mymodules.f90:33-41)."""

    _fields_ = [("lr", C.c_double * 9), ("li", C.c_double * 9), ("lv", C.c_double), ("sig", C.c_double),
                ("lammda", C.c_double), ("logerr", C.c_double), ("phi", C.c_double), ("nr", C.c_int),
                ("ni", C.c_int)]

    @property
    def LR(self):
        return np.array(self.lr[:]).reshape(3, 3, order="F")

    @property
    def LI(self):
        return np.array(self.li[:]).reshape(3, 3, order="F")


class StepConfig(C.Structure):
    """``rsd_step_config``: control settings of the periodic FTS step."""

    _fields_ = [("lb", C.c_double * 3), ("gamma", C.c_double), ("ews_err", C.c_double), ("mu_f", C.c_double),
                ("correction_method", C.c_int), ("nb", C.c_int), ("gravity", C.c_double), ("dt_dem0", C.c_double),
                ("radii_test", C.c_double), ("u_bg", C.c_double * 3), ("omega_bg", C.c_double * 3),
                ("ei_bg", C.c_double * 6)]


class Control(C.Structure):
    """``rsd_control``: control_file.dat in INIT_READ_CONTROL's order (initial_conf.f90:109-189)."""

    _fields_ = ([(k, C.c_int) for k in ("correction_method", "fts_method", "filament_solve_implicit", "uselub",
                                          "use_dem", "usecollision", "use_dem_stepper", "use_gmres", "xshear",
                                          "is_period", "simple_period", "oscillation_shear", "fix", "is_case_lattice",
                                          "beginstep", "interval", "num_threads", "timestep")] +
                [("radii_test", C.c_double), ("radii_true", C.c_double), ("g_gravity", C.c_double),
                 ("u_bg", C.c_double * 3), ("omega_bg", C.c_double * 3), ("ei_bg", C.c_double * 6)] +
                [(k, C.c_double) for k in ("mu_f", "pho_f", "re_number", "gamma_alpha", "youngs_module",
                                             "poisson_ratio", "ksp", "erest", "pho_par", "linerr")] +
                [("collision_condition", C.c_int), ("mu_calculate", C.c_int)] +
                [(k, C.c_double) for k in ("t0", "end_time", "base_dt")] + [("k_write", C.c_int)] +
                [(k, C.c_double) for k in ("lj_eps", "lj_sigma", "lj_cut")] +
                [(k, C.c_int) for k in ("boundary", "wall_method", "use_dlvo")] +
                [(k, C.c_double) for k in ("debye_length", "z0", "lambda_l", "permittivity", "potential", "ah",
                                             "surfaceenergy")] +
                [("brown", C.c_int), ("kbt", C.c_double)])


class Units(C.Structure):
    """``rsd_units``: the step's inputs in nondimensional units (mymodules.f90:808-881, main.f90:148-155)."""

    _fields_ = [(k, C.c_double) for k in ("lambda_time", "mu_f", "gravity", "end_time", "base_dt", "t0", "dt_dem0",
                                          "kbt", "frequency", "gamma")]


def _p(a):
    return a.ctypes.data_as(_dp)


def _f(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


_lib = None


def load_library(path: str | None = None):
    """Load ``librsd_amd.so`` (fails loudly if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RSDError(f"{path} not found: build it with __graft_entry__.build() "
                       f"or make -C regularized-stokesian-dynamics_amd/csrc")
    L = C.CDLL(path)
    L.rsd_last_error.restype = C.c_char_p
    L.rsd_step_config_default.argtypes = [C.POINTER(StepConfig), C.c_double]
    L.rsd_step_config_default.restype = None
    L.rsd_device_init.argtypes = [C.c_int]
    L.rsd_lub_tables_load.argtypes = [C.c_char_p]
    L.rsd_set_lattice_skew.argtypes = [C.c_double, _dp, C.c_double, C.POINTER(Lattice)]
    L.rsd_calc_lattice_inverse.argtypes = [C.POINTER(Lattice), C.c_double, C.c_int, _dp]
    L.rsd_per_skew_correction.argtypes = [C.POINTER(Lattice), _dp, _dp]
    L.rsd_per_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.POINTER(Lattice), C.c_double, C.c_int, C.c_double,
                                         _dp, _dp, _dp, _ip]
    L.rsd_per_grpery_inv_fri_images.argtypes = [C.c_int, _dp, _dp, C.POINTER(Lattice), C.c_double, C.c_int, C.c_int,
                                                C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _ip]
    L.rsd_ctx_set_ewald_images.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.rsd_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _ip]
    L.rsd_wall_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
    L.rsd_lubmxcalc.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.POINTER(Lattice), C.c_double, _dp, _dp, _dp,
                                _ip]
    L.rsd_matrev.argtypes = [C.c_int, _dp]
    L.rsd_cholesky.argtypes = [C.c_int, _dp]
    L.rsd_ctx_set_brownian.argtypes = [C.c_void_p, C.c_double, C.c_ulonglong]
    L.rsd_ctx_set_brownian_z.argtypes = [C.c_void_p, _dp]
    L.rsd_ctx_get_brownian.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.rsd_mob_to_resist_fts.argtypes = [C.c_int] + [_dp] * 6
    L.rsd_arrangement.argtypes = [C.c_int] + [_dp] * 4
    L.rsd_stepper_nogmres.argtypes = [C.c_int] + [_dp] * 11
    L.rsd_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(StepConfig)]
    L.rsd_ctx_destroy.argtypes = [C.c_void_p]
    L.rsd_ctx_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.rsd_ctx_get_state.argtypes = [C.c_void_p, _dp, _dp]
    L.rsd_ctx_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
    L.rsd_ctx_get_last.argtypes = [C.c_void_p, _dp, _dp, _ip]
    L.rsd_ctx_stage_times.argtypes = [C.c_void_p, _dp]
    L.rsd_ctx_set_substep_solver.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
    L.rsd_ctx_substep_stats.argtypes = [C.c_void_p, _ip, _ip]
    L.rsd_ctx_step_stats.argtypes = [C.c_void_p, _ip, _ip, _ip]
    L.rsd_ctx_set_resistance_route.argtypes = [C.c_void_p, C.c_int]
    L.rsd_ctx_set_geometry.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.rsd_ctx_set_iterative.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int]
    L.rsd_ctx_iterative_stats.argtypes = [C.c_void_p, _ip, _dp, _ip]
    L.rsd_ctx_set_iterative_graphs.argtypes = [C.c_void_p, C.c_int]
    L.rsd_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
    L.rsd_batch_destroy.argtypes = [C.c_void_p]
    L.rsd_batch_set_width.argtypes = [C.c_void_p, C.c_int]
    L.rsd_batch_get_width.argtypes = [C.c_void_p, _ip]
    L.rsd_batch_step.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _ip]
    L.rsd_batch_autotune.argtypes = [C.c_void_p, C.c_double, _dp, _ip, C.c_int, _ip, _dp]
    L.rsd_gnn_create.argtypes = [C.POINTER(C.c_void_p), C.c_ulonglong]
    L.rsd_gnn_destroy.argtypes = [C.c_void_p]
    L.rsd_gnn_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.rsd_gnn_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.rsd_ctx_set_gnn.argtypes = [C.c_void_p, C.c_void_p]
    L.rsd_ctx_prepare_iterative.argtypes = [C.c_void_p, C.c_double]
    L.rsd_ctx_iterative_system.argtypes = [C.c_void_p, _dp, _dp, _ip, _ip, _dp, _ip]
    L.rsd_ctx_gnn_loss.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_ulonglong, _dp, C.POINTER(C.c_float), _dp]
    L.rsd_ctx_gnn_train.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _dp]
    L.rsd_ctx_iterative_solve.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _ip, _dp, _dp]
    L.rsd_stepper_stokesian.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.POINTER(StepConfig), _dp, _dp]
    L.rsd_read_control.argtypes = [C.c_char_p, C.POINTER(Control)]
    L.rsd_read_initial_config.argtypes = [C.c_char_p, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _ip]
    L.rsd_nondimensionalise.argtypes = [C.POINTER(Control), C.POINTER(Units)]
    L.rsd_write_vtf_frame.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, C.c_double, _dp, _dp]
    L.rsd_write_yeta_mu.argtypes = [C.c_char_p, C.c_int, C.c_double, _dp, C.c_double, C.c_double, C.c_double,
                                    C.c_double]
    L.rsd_profile_enable.argtypes = [C.c_int]
    L.rsd_profile_enable_families.argtypes = [C.c_uint]
    L.rsd_profile_read.argtypes = [C.c_int, C.POINTER(C.c_long), _dp, _dp, _dp]
    L.rsd_profile_family_name.argtypes = [C.c_int]
    L.rsd_profile_family_name.restype = C.c_char_p
    _lib = L
    return L


PROFILE_FAMILIES = 11


def profile_enable(on: bool = True, families=None):
    """Start (and reset) per-kernel-family HIP-event timing: every family, or
    only those named in `families` (names as profile_read reports them)."""
    L = load_library()
    if families is None or not on:
        _check(L.rsd_profile_enable(int(on)), "rsd_profile_enable")
        return
    names = [L.rsd_profile_family_name(f).decode() for f in range(PROFILE_FAMILIES)]
    mask = 0
    for name in families:
        if name not in names:
            raise ValueError(f"unknown kernel family {name!r}")
        mask |= 1 << names.index(name)
    _check(L.rsd_profile_enable_families(mask), "rsd_profile_enable_families")


def profile_read():
    """{family name: (launches, total ms, algorithmic flops, bytes)} since profile_enable."""
    L = load_library()
    out = {}
    for f in range(PROFILE_FAMILIES):
        n, ms, fl, by = C.c_long(), np.zeros(1), np.zeros(1), np.zeros(1)
        _check(L.rsd_profile_read(f, C.byref(n), _p(ms), _p(fl), _p(by)), "rsd_profile_read")
        out[L.rsd_profile_family_name(f).decode()] = (n.value, ms[0], fl[0], by[0])
    return out


def _check(rc, what):
    if rc < 0:
        msg = _lib.rsd_last_error().decode(errors="replace")
        raise RSDError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")
    return rc


def init(device: int = 0, lubdat: str | None = None):
    """Select the device and load the lubrication tables (hydro_lub.f90:529-596)."""
    L = load_library()
    _check(L.rsd_device_init(device), "rsd_device_init")
    _check(L.rsd_lub_tables_load((lubdat or LUBDAT_DIR).encode()), "rsd_lub_tables_load")
    return L


def default_config(box_edge: float, **kw) -> StepConfig:
    L = load_library()
    cfg = StepConfig()
    L.rsd_step_config_default(C.byref(cfg), float(box_edge))
    for k, v in kw.items():
        if k in ("lb", "u_bg", "omega_bg", "ei_bg"):
            getattr(cfg, k)[:] = list(v)
        else:
            setattr(cfg, k, v)
    return cfg


# ---------------------------------------------------------- on-disk formats
def read_control(path) -> Control:
    """INIT_READ_CONTROL (initial_conf.f90:109-189) of a control_file.dat."""
    c = Control()
    _check(load_library().rsd_read_control(os.fsencode(path), C.byref(c)), "read_control")
    return c


def read_initial_config(path):
    """INITIAL_SIZE + INITIAL_CONF (initial_conf.f90:356-476):
    -> (conf (N,3), radii, LB, T, u_pos (6N; translations of a hot start), hot_start)."""
    L = load_library()
    n, hot = C.c_int(), C.c_int()
    lb, t = np.zeros(3), np.zeros(1)
    _check(L.rsd_read_initial_config(os.fsencode(path), 0, C.byref(n), _p(lb), None, None, None, None,
                                     C.byref(hot)), "read_initial_config")
    N = n.value
    conf, radii, up = np.zeros((N, 3)), np.zeros(N), np.zeros(6 * N)
    _check(L.rsd_read_initial_config(os.fsencode(path), N, C.byref(n), _p(lb), _p(t), _p(conf), _p(radii), _p(up),
                                     C.byref(hot)), "read_initial_config")
    return conf, radii, lb, float(t[0]), up, bool(hot.value)


def nondimensionalise(ctl: Control) -> Units:
    u = Units()
    _check(load_library().rsd_nondimensionalise(C.byref(ctl), C.byref(u)), "nondimensionalise")
    return u


def write_vtf_frame(path, conf, t, radii, u_pos, append=True, nb=0):
    """VMD_WRITE (initial_conf.f90:250-262)."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    u_pos = np.ascontiguousarray(u_pos, dtype=np.float64)
    _check(load_library().rsd_write_vtf_frame(os.fsencode(path), int(append), len(radii), int(nb), _p(conf),
                                              float(t), _p(radii), _p(u_pos)), "write_vtf_frame")


def write_yeta_mu(path, t, yeta_mu, lb1, frequency, gamma, gamma_angle=0.0, append=True):
    """The yeta_mu.txt row of main.f90:197-201."""
    y = np.ascontiguousarray(yeta_mu, dtype=np.float64)
    _check(load_library().rsd_write_yeta_mu(os.fsencode(path), int(append), float(t), _p(y), float(lb1),
                                            float(frequency), float(gamma), float(gamma_angle)), "write_yeta_mu")


RUN_PATH = os.path.join(PKG_DIR, "rsd_run")


# ---------------------------------------------------------------- lattice
def set_lattice_skew(t, lb, gamma) -> Lattice:
    """SET_LATTICE_SKEW (csd_post.f90:105-136)."""
    lat = Lattice()
    lb = np.ascontiguousarray(lb, dtype=np.float64)
    _check(load_library().rsd_set_lattice_skew(float(t), _p(lb), float(gamma), C.byref(lat)), "set_lattice_skew")
    return lat


def calc_lattice_inverse(lat: Lattice, ews_err, radii) -> Lattice:
    """CALC_LATTICE_INVERSE (csd_post.f90:158-183)."""
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    _check(load_library().rsd_calc_lattice_inverse(C.byref(lat), float(ews_err), len(radii), _p(radii)),
           "calc_lattice_inverse")
    return lat


def per_skew_correction(lat: Lattice, r):
    """PER_SKEW_CORRECTION (csd_post.f90:24-63)."""
    r = np.array(r, dtype=np.float64)
    d = np.zeros(1)
    _check(load_library().rsd_per_skew_correction(C.byref(lat), _p(r), _p(d)), "per_skew_correction")
    return r, d[0]


# ---------------------------------------------------------------- drop-ins
def per_grpery_inv_fri(conf, radii, lat: Lattice, ews_err, correction_method=2, mu_f=1.0 / PAI):
    """PER_GRPERY_INV_FRI (per_GRPY_YZD.f90:13-200) -> APP, APQ, AQQ, n_overlap."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    N = len(radii)
    APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
    nov = C.c_int(0)
    _check(load_library().rsd_per_grpery_inv_fri(N, _p(conf), _p(radii), C.byref(lat), float(ews_err),
                                                 int(correction_method), float(mu_f), _p(APP), _p(APQ), _p(AQQ),
                                                 C.byref(nov)), "per_grpery_inv_fri")
    return APP, APQ, AQQ, nov.value


def per_grpery_inv_fri_images(conf, radii, lat: Lattice, ews_err, nr, ni, kmax2=0, correction_method=2,
                              mu_f=1.0 / PAI):
    """PER_GRPERY_INV_FRI with an explicit image set (real cube nr, reciprocal
    cube ni, spherical cutoff |n|^2 <= kmax2 when > 0) -> APP, APQ, AQQ, n_overlap."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    N = len(radii)
    APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
    nov = C.c_int(0)
    _check(load_library().rsd_per_grpery_inv_fri_images(N, _p(conf), _p(radii), C.byref(lat), float(ews_err), int(nr),
                                                        int(ni), int(kmax2), int(correction_method), float(mu_f),
                                                        _p(APP), _p(APQ), _p(AQQ), C.byref(nov)),
           "per_grpery_inv_fri_images")
    return APP, APQ, AQQ, nov.value


def grpery_inv_fri(conf, radii, correction_method=2, mu_f=1.0 / PAI):
    """GRPERY_INV_FRI (GRPY_YZD.f90:14-264), free space -> APP, APQ, AQQ, n_overlap."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    N = len(radii)
    APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
    nov = C.c_int(0)
    _check(load_library().rsd_grpery_inv_fri(N, _p(conf), _p(radii), int(correction_method), float(mu_f), _p(APP),
                                             _p(APQ), _p(AQQ), C.byref(nov)), "grpery_inv_fri")
    return APP, APQ, AQQ, nov.value


def wall_grpery_inv_fri(conf, radii, mu_f=1.0 / PAI):
    """Wall_GRPERY_INV_FRI (Wall_interaction.f90:29-149) -> TPP (3N x 3N)."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    N = len(radii)
    T = _f((3 * N, 3 * N))
    _check(load_library().rsd_wall_grpery_inv_fri(N, _p(conf), _p(radii), float(mu_f), _p(T)), "wall_grpery_inv_fri")
    return T


def lubmxcalc(conf, radii, lat: Lattice, mu_f=1.0 / PAI, nb=0, is_period=True):
    """lubmxcalc (hydro_lub.f90:22-308) -> rfu, rfe, rse, n_contact."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    N = len(radii)
    rfu, rfe, rse = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
    nc = C.c_int(0)
    _check(load_library().rsd_lubmxcalc(N, int(nb), int(is_period), _p(conf), _p(radii), C.byref(lat), float(mu_f),
                                        _p(rfu), _p(rfe), _p(rse), C.byref(nc)), "lubmxcalc")
    return rfu, rfe, rse, nc.value


def matrev(A):
    """MATREV (matrices.f90:4-34): SPD inverse from the upper triangle -> (inverse, INFO)."""
    A = np.array(A, dtype=np.float64, order="F")
    info = _check(load_library().rsd_matrev(A.shape[0], _p(A)), "matrev")
    return A, info


def cholesky(A):
    """CHOLESKY (matrices.f90:126-157): lower factor, strict upper zeroed -> (L, INFO)."""
    A = np.array(A, dtype=np.float64, order="F")
    info = _check(load_library().rsd_cholesky(A.shape[0], _p(A)), "cholesky")
    return A, info


def mob_to_resist_fts(APP, APQ, AQQ):
    """MobToResist_FTS (matrices.f90:393-416) -> RPP, RPQ, RQQ, INFO."""
    N = APP.shape[0] // 6
    out = [_f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))]
    ins = [np.asfortranarray(x, dtype=np.float64) for x in (APP, APQ, AQQ)]
    info = _check(load_library().rsd_mob_to_resist_fts(N, *[_p(x) for x in ins], *[_p(x) for x in out]),
                  "mob_to_resist_fts")
    return (*out, info)


def arrangement(RPP, RPQ, RQQ):
    """arrangement (matrices.f90:432-445) -> mu = RPQ^T inv(RPP) RPQ - RQQ."""
    N = RPP.shape[0] // 6
    mu = _f((5 * N, 5 * N))
    ins = [np.asfortranarray(x, dtype=np.float64) for x in (RPP, RPQ, RQQ)]
    _check(load_library().rsd_arrangement(N, *[_p(x) for x in ins], _p(mu)), "arrangement")
    return mu


def stepper_nogmres(APP, APQ, AQQ, rfu, rfe, rse, Fe, EIN):
    """STEPPER_NOGmres (stepper.f90:348-410) -> U_par, SijN_hyd, RPP."""
    N = APP.shape[0] // 6
    ins = [np.asfortranarray(x, dtype=np.float64) for x in (APP, APQ, AQQ, rfu, rfe, rse)]
    Fe = np.ascontiguousarray(Fe, dtype=np.float64)
    EIN = np.ascontiguousarray(EIN, dtype=np.float64)
    u, s, rpp = np.zeros(6 * N), np.zeros(5 * N), _f((6 * N, 6 * N))
    _check(load_library().rsd_stepper_nogmres(N, *[_p(x) for x in ins], _p(Fe), _p(EIN), _p(u), _p(s), _p(rpp)),
           "stepper_nogmres")
    return u, s, rpp


# ---------------------------------------------------------------- context
GNN_PARAMS = 50629


class Gnn:
    """The learned preconditioner's model (``rsd_gnn``): a graph network whose
    FP32 parameters live on the device."""

    def __init__(self, seed=0, params=None):
        load_library()
        self._g = C.c_void_p()
        _check(_lib.rsd_gnn_create(C.byref(self._g), int(seed)), "rsd_gnn_create")
        if params is not None:
            self.set_params(params)

    def params(self):
        p = np.zeros(GNN_PARAMS, np.float32)
        _check(_lib.rsd_gnn_get_params(self._g, p.ctypes.data_as(C.POINTER(C.c_float))), "rsd_gnn_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (GNN_PARAMS,):
            raise ValueError(f"expected {GNN_PARAMS} parameters, got {p.shape}")
        _check(_lib.rsd_gnn_set_params(self._g, p.ctypes.data_as(C.POINTER(C.c_float))), "rsd_gnn_set_params")

    def save(self, path):
        np.save(path, self.params())

    @classmethod
    def load(cls, path):
        return cls(params=np.load(path, allow_pickle=False))

    def close(self):
        if self._g:
            _lib.rsd_gnn_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent suspensions (Suspension objects) stepped together
    (``rsd_batch``): concurrent host threads, one context each at a time."""

    def __init__(self, suspensions, width=None):
        load_library()
        self.sus = list(suspensions)
        arr = (C.c_void_p * len(self.sus))(*[s._ctx.value for s in self.sus])
        self._b = C.c_void_p()
        _check(_lib.rsd_batch_create(C.byref(self._b), len(self.sus), arr), "rsd_batch_create")
        if width is not None:
            self.set_width(width)

    def set_width(self, width):
        _check(_lib.rsd_batch_set_width(self._b, int(width)), "rsd_batch_set_width")

    def width(self):
        w = C.c_int()
        _check(_lib.rsd_batch_get_width(self._b, C.byref(w)), "rsd_batch_get_width")
        return w.value

    def step(self, dt, t):
        """One step of every suspension at its own time t[i]; returns
        (yeta_mu (count, 5), INFO per suspension)."""
        n = len(self.sus)
        tt = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
        y = np.zeros((n, 5))
        info = np.zeros(n, np.int32)
        _check(_lib.rsd_batch_step(self._b, float(dt), _p(tt), _p(y), info.ctypes.data_as(_ip)), "rsd_batch_step")
        return y, info

    def autotune(self, dt, t, widths=(1, 2, 4, 8, 16, 32)):
        """Time one batch step per candidate width (from the same states, which
        are restored) and keep the fastest; returns (best width, {width: ms})."""
        n = len(self.sus)
        tt = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
        cand = np.ascontiguousarray(widths, dtype=np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int()
        _check(_lib.rsd_batch_autotune(self._b, float(dt), _p(tt), cand.ctypes.data_as(_ip), len(cand),
                                       C.byref(best), _p(ms)), "rsd_batch_autotune")
        return best.value, {int(w): float(m) for w, m in zip(cand, ms)}

    def close(self):
        if self._b:
            _lib.rsd_batch_destroy(self._b)
            self._b = C.c_void_p()


class Suspension:
    """Device-resident suspension: one ``rsd_ctx`` (STEPPER_Stokesian state)."""

    def __init__(self, conf, radii, cfg: StepConfig, u_pos=None):
        L = load_library()
        self.N = len(radii)
        self.cfg = cfg
        self._ctx = C.c_void_p()
        _check(L.rsd_ctx_create(C.byref(self._ctx), self.N, C.byref(cfg)), "rsd_ctx_create")
        self.set_state(conf, radii, u_pos)

    def set_state(self, conf, radii, u_pos=None):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        self.radii = np.ascontiguousarray(radii, dtype=np.float64)
        up = None if u_pos is None else _p(np.ascontiguousarray(u_pos, dtype=np.float64))
        _check(_lib.rsd_ctx_set_state(self._ctx, _p(conf), _p(self.radii), up), "rsd_ctx_set_state")

    def step(self, dt, t):
        """STEPPER_Stokesian (stepper.f90:33-156); returns (yeta_mu[5], INFO)."""
        y = np.zeros(5)
        info = _check(_lib.rsd_ctx_step(self._ctx, float(dt), float(t), _p(y)), "rsd_ctx_step")
        return y, info

    def state(self):
        conf, up = np.zeros((self.N, 3)), np.zeros(6 * self.N)
        _check(_lib.rsd_ctx_get_state(self._ctx, _p(conf), _p(up)), "rsd_ctx_get_state")
        return conf, up

    def last(self):
        u, s, nt = np.zeros(6 * self.N), np.zeros(5 * self.N), C.c_int()
        _check(_lib.rsd_ctx_get_last(self._ctx, _p(u), _p(s), C.byref(nt)), "rsd_ctx_get_last")
        return u, s, nt.value

    SUBSTEP_DIRECT = 0
    SUBSTEP_PCG = 1
    SUBSTEP_PCG_ONE_LAUNCH = 2

    def set_substep_solver(self, mode, tol=1e-14, max_iter=60):
        """Solver for the later collision sub-steps: SUBSTEP_DIRECT factors
        RPP at every sub-step (the reference's arithmetic); SUBSTEP_PCG (the
        default) runs preconditioned CG from the first sub-step's inverse, one
        launch per CG phase, and redoes the step's sub-steps with the
        factorisation if a solve has not converged after max_iter iterations;
        SUBSTEP_PCG_ONE_LAUNCH runs each CG solve as one cooperative launch."""
        _check(_lib.rsd_ctx_set_substep_solver(self._ctx, int(mode), float(tol), int(max_iter)),
               "rsd_ctx_set_substep_solver")

    RESIST_EXPLICIT = 0
    RESIST_SCHUR = 1
    RESIST_ITERATIVE = 2
    PRECOND_NONE = 0
    PRECOND_BLOCK_JACOBI = 1
    PRECOND_MIXED = 2
    PRECOND_LEARNED = 3

    def set_resistance_route(self, route):
        """RESIST_EXPLICIT inverts the grand mobility whole (MobToResist_FTS);
        RESIST_SCHUR (the default) inverts its Schur complement for W_PP and
        forms the other blocks only as the products the step needs;
        RESIST_ITERATIVE solves the mobility form of each sub-step by GMRES
        (STEPPER_Gmres, stepper.f90:450-546) without inverting anything."""
        _check(_lib.rsd_ctx_set_resistance_route(self._ctx, int(route)), "rsd_ctx_set_resistance_route")

    def set_iterative(self, precond=2, tol=1e-12, max_iter=400, inner_iters=20):
        """The iterative route's preconditioner (PRECOND_*), relative residual
        tolerance, GMRES iteration cap per solve and, for PRECOND_MIXED, the
        FP32 inner iterations."""
        _check(_lib.rsd_ctx_set_iterative(self._ctx, int(precond), float(tol), int(max_iter), int(inner_iters)),
               "rsd_ctx_set_iterative")

    PRECOND_MIXED_LEARNED = 4

    def set_gnn(self, gnn):
        """Attach a learned preconditioner model (Gnn) for PRECOND_LEARNED and
        PRECOND_MIXED_LEARNED (None detaches)."""
        self._gnn = gnn
        _check(_lib.rsd_ctx_set_gnn(self._ctx, gnn._g if gnn is not None else None), "rsd_ctx_set_gnn")

    def prepare_iterative(self, t=0.0):
        """Assemble the iterative route's system at the current positions."""
        _check(_lib.rsd_ctx_prepare_iterative(self._ctx, float(t)), "rsd_ctx_prepare_iterative")

    def iterative_system(self, dense=True):
        """(M, R_lub, nnb, col, dist, cap) of the prepared system (M and R as
        11N x 11N arrays when dense)."""
        N = self.N
        cap = C.c_int()
        _check(_lib.rsd_ctx_iterative_system(self._ctx, None, None, None, None, None, C.byref(cap)),
               "rsd_ctx_iterative_system")
        cap = cap.value
        M = _f((11 * N, 11 * N)) if dense else None
        R = _f((11 * N, 11 * N)) if dense else None
        nnb, col = np.zeros(N, np.int32), np.zeros(N * cap, np.int32)
        dist = np.zeros(N * cap)
        _check(_lib.rsd_ctx_iterative_system(self._ctx, _p(M) if dense else None, _p(R) if dense else None,
                                              nnb.ctypes.data_as(_ip), col.ctypes.data_as(_ip), _p(dist), None),
               "rsd_ctx_iterative_system")
        return M, R, nnb, col.reshape(N, cap), dist.reshape(N, cap), cap

    def gnn_loss(self, gnn, nz=4, seed=0, grad=True):
        """(loss, gradient or None, z) of the prepared system for the model."""
        loss = C.c_double()
        g = np.zeros(GNN_PARAMS, np.float32) if grad else None
        z = np.zeros((nz, 11 * self.N))
        _check(_lib.rsd_ctx_gnn_loss(self._ctx, gnn._g, int(nz), int(seed), C.byref(loss),
                                     g.ctypes.data_as(C.POINTER(C.c_float)) if grad else None, _p(z)),
               "rsd_ctx_gnn_loss")
        return loss.value, g, z

    def gnn_train(self, gnn, iters, nz=4, lr=1e-3, seed=0):
        """iters Adam steps on the prepared system; returns the losses."""
        losses = np.zeros(int(iters))
        _check(_lib.rsd_ctx_gnn_train(self._ctx, gnn._g, int(iters), int(nz), float(lr), int(seed), _p(losses)),
               "rsd_ctx_gnn_train")
        return losses

    def iterative_solve(self, precond, tol=1e-10, max_iter=4000):
        """One solve of the prepared system from zero: (iterations, relres, ms)."""
        it, rr, ms = C.c_int(), C.c_double(), C.c_double()
        _check(_lib.rsd_ctx_iterative_solve(self._ctx, int(precond), float(tol), int(max_iter), C.byref(it),
                                            C.byref(rr), C.byref(ms)), "rsd_ctx_iterative_solve")
        return it.value, rr.value, ms.value

    def set_iterative_graphs(self, on=True):
        """Replay the mixed preconditioners' inner GMRES as a captured HIP graph
        (the default) or launch it kernel by kernel."""
        _check(_lib.rsd_ctx_set_iterative_graphs(self._ctx, int(bool(on))), "rsd_ctx_set_iterative_graphs")

    def iterative_stats(self):
        """(GMRES iterations of the last step, largest final relative
        residual, whether a solve failed and the step was redone directly)."""
        it, fb = C.c_int(), C.c_int()
        rr = C.c_double()
        _check(_lib.rsd_ctx_iterative_stats(self._ctx, C.byref(it), C.byref(rr), C.byref(fb)),
               "rsd_ctx_iterative_stats")
        return it.value, rr.value, bool(fb.value)

    def set_geometry(self, is_period=True, wall=False):
        """IsPeriod = simplePeriod (Ewald under Lees-Edwards images, or free
        space) and Wall_method (stepper.f90:54-70)."""
        _check(_lib.rsd_ctx_set_geometry(self._ctx, int(is_period), int(wall)), "rsd_ctx_set_geometry")

    def set_ewald_images(self, nr=-1, ni=-1, kmax2=0):
        """Explicit Ewald image set for the step (nr < 0: the reference's rule)."""
        _check(_lib.rsd_ctx_set_ewald_images(self._ctx, int(nr), int(ni), int(kmax2)), "rsd_ctx_set_ewald_images")

    def set_brownian(self, kbt, seed=0):
        """BROWN with this kBT (0: off); z from set_brownian_z, else the device generator keyed by seed."""
        _check(_lib.rsd_ctx_set_brownian(self._ctx, float(kbt), int(seed)), "rsd_ctx_set_brownian")

    def set_brownian_z(self, z):
        """The next step's 3N standard normals."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        assert z.shape == (3 * self.N,)
        _check(_lib.rsd_ctx_set_brownian_z(self._ctx, _p(z)), "rsd_ctx_set_brownian_z")

    def brownian(self):
        """(z, u_brown, SijN_B) of the last step."""
        z, ub, sb = np.zeros(3 * self.N), np.zeros(3 * self.N), np.zeros(5 * self.N)
        _check(_lib.rsd_ctx_get_brownian(self._ctx, _p(z), _p(ub), _p(sb)), "rsd_ctx_get_brownian")
        return z, ub, sb

    def substep_stats(self):
        """(CG iterations, sub-steps solved directly after CG) of the last step."""
        it, fb = C.c_int(), C.c_int()
        _check(_lib.rsd_ctx_substep_stats(self._ctx, C.byref(it), C.byref(fb)), "rsd_ctx_substep_stats")
        return it.value, fb.value

    def step_stats(self):
        """(CG iterations, CG solves that fell back to the factorisation,
        1 if the step was redone through the explicit inverse) of the last step."""
        it, fb, rd = C.c_int(), C.c_int(), C.c_int()
        _check(_lib.rsd_ctx_step_stats(self._ctx, C.byref(it), C.byref(fb), C.byref(rd)), "rsd_ctx_step_stats")
        return it.value, fb.value, rd.value

    def last_nt(self):
        """NT_DEM of the last step (no device copies)."""
        nt = C.c_int()
        _check(_lib.rsd_ctx_get_last(self._ctx, None, None, C.byref(nt)), "rsd_ctx_get_last")
        return nt.value

    def stage_times(self):
        ms = np.zeros(6)
        _check(_lib.rsd_ctx_stage_times(self._ctx, _p(ms)), "rsd_ctx_stage_times")
        return ms

    def close(self):
        if getattr(self, "_ctx", None):
            _lib.rsd_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
