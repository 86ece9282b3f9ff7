"""Python face of the oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product path never does.

Two checkers live here:

* :class:`Oracle` wraps ``librsd_oracle.so`` (the C restatement of the
  reference, ``rsd_oracle.c``) and adds the dense algebra of
  ``matrices.f90`` / ``stepper.f90`` in numpy + LAPACK (scipy), exactly the
  LAPACK routines the reference calls (DPOTRF/DPOTRI, ``matrices.f90:4-34``).
* :class:`Reference` wraps ``oracle/_ref/libref.so``: the reference's own
  Fortran, compiled with flang by ``oracle/Makefile``, driven through
  ``ref_driver.f90``.  It pins the restatement.

Array conventions follow the reference: positions ``conf`` are numpy arrays
of shape (N, 3) in C order (the memory of Fortran ``CONF(3,N)``); matrices
are Fortran-ordered (column-major) float64 arrays.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg.lapack as lapack

HERE = os.path.dirname(os.path.abspath(__file__))
PAI = np.pi

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a):
    return a.ctypes.data_as(_dp)


class Lattice(C.Structure):
    """Mirror of ``orc_lattice`` (the reference's LATTICE_BASE module state)."""

    _fields_ = [("LR", C.c_double * 9), ("LI", C.c_double * 9), ("LV", C.c_double),
                ("sig", C.c_double), ("lam", C.c_double), ("logerr", C.c_double),
                ("phi", C.c_double), ("NR", C.c_int), ("NI", C.c_int)]

    @property
    def lr(self):
        return np.array(self.LR[:]).reshape(3, 3, order="F")

    @property
    def li(self):
        return np.array(self.LI[:]).reshape(3, 3, order="F")


def _f(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


_BIG_STACK = 8 << 30


def _on_big_stack(fn, *args, **kw):
    """Run fn in a thread with an 8 GiB stack (reserved, not committed).  The flang-built reference keeps
    its automatic arrays (grmobmx, Mtemp, APPl: tens of MB at N>=64) on the
    calling thread's stack, which overflows the default 8 MiB."""
    import threading
    box = {}

    def run():
        try:
            box["r"] = fn(*args, **kw)
        except BaseException as e:  # re-raised in the caller
            box["e"] = e

    old = threading.stack_size(_BIG_STACK)
    try:
        t = threading.Thread(target=run)
        t.start()
    finally:
        threading.stack_size(old)
    t.join()
    if "e" in box:
        raise box["e"]
    return box.get("r")


@contextlib.contextmanager
def _quiet_stdout(flush=None):
    """The reference prints progress on every call; silence fd 1 around it.
    ``flush`` empties the Fortran runtime's own buffer before fd 1 returns."""
    import sys
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        yield
    finally:
        if flush is not None:
            flush()
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


# --------------------------------------------------------------------- oracle
class Oracle:
    """C restatement + numpy/LAPACK dense algebra."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "librsd_oracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        L.orc_init()
        L.orc_set_lattice_skew.argtypes = [C.c_double, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_calc_lattice_inverse.argtypes = [C.POINTER(Lattice), C.c_double, C.c_int, _dp]
        L.orc_per_skew_correction.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_period_correction.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.orc_rp_self.argtypes = [C.POINTER(Lattice), C.c_double, _dp, _dp, _dp]
        six = [_dp] * 6
        L.orc_rp_ij.argtypes = [C.POINTER(Lattice), _dp, C.c_double, C.c_double] + six
        L.orc_rp_regular.argtypes = [_dp, C.c_double, C.c_double] + six
        L.orc_rp_free_fts.argtypes = [_dp, C.c_double, C.c_double] + six
        L.orc_overlap_regular_correction.argtypes = [_dp, C.c_double, C.c_double] + six
        L.orc_rp_yamakawa.argtypes = [_dp, C.c_double, C.c_double] + six
        L.orc_overlap_yamakawa_correction.argtypes = [_dp, C.c_double, C.c_double] + six
        L.orc_per_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.POINTER(Lattice), C.c_double, C.c_int,
                                             C.c_double, _dp, _dp, _dp, _ip]
        L.orc_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _ip]
        L.orc_wall_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
        L.orc_wall_grpery_inv_fri.restype = None
        L.orc_set_images.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_set_images.restype = None
        L.orc_lub_tables_size.restype = C.c_int
        self._tab = C.create_string_buffer(L.orc_lub_tables_size())
        L.orc_lub_tables_read.argtypes = [C.c_char_p, C.c_void_p]
        L.orc_lubcalc.argtypes = [C.c_double, C.c_void_p, _dp]
        L.orc_lubmxcalc.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_void_p,
                                    _dp, _dp, _dp]
        L.orc_calc_uo_bg.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_ei_transy2.argtypes = [_dp, _dp]
        L.orc_pos_collision.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip]
        self.tables_loaded = False

    # ---- lattice (csd_post.f90 period_bdy_tools)
    def lattice(self, T, LB, gamma):
        lat = Lattice()
        LB = np.ascontiguousarray(LB, dtype=np.float64)
        lr, li, lv = np.zeros(9), np.zeros(9), np.zeros(1)
        self.lib.orc_set_lattice_skew(T, _p(LB), gamma, _p(lr), _p(li), _p(lv))
        lat.LR[:] = lr.tolist()
        lat.LI[:] = li.tolist()
        lat.LV = lv[0]
        return lat

    def calc_lattice_inverse(self, lat, ews_err, radii):
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        self.lib.orc_calc_lattice_inverse(C.byref(lat), ews_err, len(radii), _p(radii))
        return lat

    def per_skew_correction(self, lat, R):
        R = np.array(R, dtype=np.float64)
        d = np.zeros(1)
        lr, li = np.array(lat.LR[:]), np.array(lat.LI[:])
        self.lib.orc_per_skew_correction(_p(lr), _p(li), _p(R), _p(d))
        return R, d[0]

    def period_correction(self, R, LB, T, gamma):
        R = np.array(R, dtype=np.float64)
        LB = np.ascontiguousarray(LB, dtype=np.float64)
        self.lib.orc_period_correction(_p(R), _p(LB), T, gamma)
        return R

    # ---- pair tensors (Ewald_summation.f90, Regularization.f90, stokesian_Green.f90)
    def rp_self(self, lat, a):
        tt, rr, qq = _f((3, 3)), _f((3, 3)), _f((5, 5))
        self.lib.orc_rp_self(C.byref(lat), a, _p(tt), _p(rr), _p(qq))
        return tt, rr, qq

    def _six(self, fn, *args):
        o = [_f((3, 3)), _f((3, 3)), _f((3, 3)), _f((5, 5)), _f((3, 5)), _f((3, 5))]
        fn(*args, *[_p(x) for x in o])
        return o  # TT, RR, TR, QQ, GQ, HQ

    def rp_ij(self, lat, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        return self._six(self.lib.orc_rp_ij, C.byref(lat), _p(R), ai, aj)

    def rp_regular(self, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        return self._six(self.lib.orc_rp_regular, _p(R), ai, aj)

    def rp_free_fts(self, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        return self._six(self.lib.orc_rp_free_fts, _p(R), ai, aj)

    def overlap_correction(self, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        return self._six(self.lib.orc_overlap_regular_correction, _p(R), ai, aj)

    def rp_yamakawa(self, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        return self._six(self.lib.orc_rp_yamakawa, _p(R), ai, aj)

    def overlap_yamakawa_correction(self, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        return self._six(self.lib.orc_overlap_yamakawa_correction, _p(R), ai, aj)

    def set_images(self, nr=-1, ni=-1, kmax2=0):
        """Explicit Ewald image sets for the grand mobility (nr < 0: the reference rule)."""
        self.lib.orc_set_images(int(nr), int(ni), int(kmax2))

    # ---- grand mobility (per_GRPY_YZD.f90:13-200)
    def per_grpery_inv_fri(self, conf, radii, lat, ews_err, correction_method=2, mu_f=1.0 / PAI):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
        nov = C.c_int(0)
        rc = self.lib.orc_per_grpery_inv_fri(N, _p(conf), _p(radii), C.byref(lat), ews_err, correction_method,
                                             mu_f, _p(APP), _p(APQ), _p(AQQ), C.byref(nov))
        if rc != 0:
            raise RuntimeError(f"orc_per_grpery_inv_fri failed: {rc}")
        return APP, APQ, AQQ, nov.value

    # ---- free space and wall (GRPY_YZD.f90:14-264, Wall_interaction.f90:29-149)
    def grpery_inv_fri(self, conf, radii, correction_method=2, mu_f=1.0 / PAI):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
        nov = C.c_int(0)
        rc = self.lib.orc_grpery_inv_fri(N, _p(conf), _p(radii), correction_method, mu_f, _p(APP), _p(APQ), _p(AQQ),
                                         C.byref(nov))
        if rc != 0:
            raise RuntimeError(f"orc_grpery_inv_fri failed: {rc}")
        return APP, APQ, AQQ, nov.value

    def wall_grpery_inv_fri(self, conf, radii, mu_f=1.0 / PAI):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        T = _f((3 * N, 3 * N))
        self.lib.orc_wall_grpery_inv_fri(N, _p(conf), _p(radii), mu_f, _p(T))
        return T

    # ---- lubrication (hydro_lub.f90)
    def load_tables(self, lubdat_dir):
        rc = self.lib.orc_lub_tables_read(lubdat_dir.encode(), self._tab)
        if rc != 0:
            raise FileNotFoundError(f"lubrication tables not found under {lubdat_dir}")
        self.tables_loaded = True

    def lubcalc(self, s):
        o = np.zeros(19)
        self.lib.orc_lubcalc(s, self._tab, _p(o))
        return o

    def lubmxcalc(self, conf, radii, lat, mu_f=1.0 / PAI, nb=0, is_period=True):
        assert self.tables_loaded
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        rfu, rfe, rse = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
        lr, li = np.array(lat.LR[:]), np.array(lat.LI[:])
        nc = self.lib.orc_lubmxcalc(N, nb, int(is_period), _p(conf), _p(radii), _p(lr), _p(li), mu_f, self._tab,
                                    _p(rfu), _p(rfe), _p(rse))
        return rfu, rfe, rse, nc

    def ei_transy2(self, E):
        E = np.asfortranarray(E, dtype=np.float64)
        o = np.zeros(5)
        self.lib.orc_ei_transy2(_p(o), _p(E))
        return o

    def calc_uo_bg(self, conf, u_bg, omega_bg, Eij):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        N = conf.shape[0]
        out = np.zeros(6 * N)
        ub = np.ascontiguousarray(u_bg, dtype=np.float64)
        ob = np.ascontiguousarray(omega_bg, dtype=np.float64)
        E = np.asfortranarray(Eij, dtype=np.float64)
        self.lib.orc_calc_uo_bg(N, _p(conf), _p(ub), _p(ob), _p(E), _p(out))
        return out

    def pos_collision(self, conf, radii, nb=0):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        floc = np.ones(N, dtype=np.int32)
        if nb:
            floc[N - nb:] = 0
        return bool(self.lib.orc_pos_collision(N, nb, _p(conf), _p(radii), floc.ctypes.data_as(_ip)))

    # ---- dense algebra (matrices.f90 / stepper.f90), LAPACK as in the reference
    @staticmethod
    def matrev(A):
        """matrices.f90:4-34 MATREV: DPOTRF('U') + DPOTRI('U') + mirror upper->lower.
        Returns (inverse, INFO of DPOTRF)."""
        c, info = lapack.dpotrf(np.asfortranarray(A), lower=0, clean=0)
        inv, info2 = lapack.dpotri(c, lower=0)
        inv = np.triu(inv) + np.triu(inv, 1).T
        return np.asfortranarray(inv), int(info)

    def mob_to_resist_fts(self, APP, APQ, AQQ):
        """matrices.f90:393-416 MobToResist_FTS."""
        n6 = APP.shape[0]
        M = np.block([[APP, APQ], [APQ.T, AQQ]])
        Mi, info = self.matrev(M)
        return Mi[:n6, :n6].copy(order="F"), Mi[:n6, n6:].copy(order="F"), Mi[n6:, n6:].copy(order="F"), info

    def arrangement(self, RPP, RPQ, RQQ):
        """matrices.f90:432-445: mu = RPQ^T inv(RPP) RPQ - RQQ."""
        Ri, _ = self.matrev(RPP)
        return RPQ.T @ Ri @ RPQ - RQQ

    def stepper_nogmres(self, APP, APQ, AQQ, rfu, rfe, rse, Fe, EIN):
        """stepper.f90:348-410 (u_brown = 0, Fe_inter = 0 as STEPPER_Stokesian passes them).
        Returns U_par (6N), SijN_hyd (5N), RPP."""
        RPP0, RPQ0, RQQ0, _ = self.mob_to_resist_fts(APP, APQ, AQQ)
        RPP, RPQ, RQQ = RPP0 + rfu, RPQ0 + rfe, RQQ0 + rse          # arrangResist matrices.f90:218-228
        APPR, _ = self.matrev(RPP)
        U_par = APPR @ (Fe + RPQ @ EIN)
        mu = self.arrangement(RPP, RPQ, RQQ)
        SijN_hyd = mu @ EIN
        self.last_RPQ = RPQ          # SijN_B = RPQ^T UB (stepper.f90:407) needs it
        return U_par, SijN_hyd, RPP

    # ---- the iterative formulation (hydro_SD_FTS, stepper.f90:476-546)
    @staticmethod
    def mobility_form(APP, APQ, AQQ, rfu, rfe, rse):
        """The grand mobility M and lubrication R_lub (11N, field-major) and
        K = [(I + R_lub M)_P ; M_Q], whose solution of K x = [F; -EIN] gives
        U = (M x)_P and SijN = (x + R_lub M x)_Q: the single system behind
        hydro_SD_FTS's pair of alternating GMRES solves (grmb_LHS/RHS and
        RM_LHS/RHS, stepper.f90:492-520), which the reference stops once
        ||[Fe; -EIN] - eval|| < LINERR 1e-2 (stepper.f90:523-530)."""
        n6 = APP.shape[0]
        M = np.block([[APP, APQ], [APQ.T, AQQ]])
        R = np.block([[rfu, rfe], [rfe.T, rse]])
        RM = R @ M
        K = M.copy()
        K[:n6] = RM[:n6]
        K[:n6, :n6] += np.eye(n6)
        return M, R, K

    def stepper_mobility_form(self, APP, APQ, AQQ, rfu, rfe, rse, Fe, EIN):
        """U_par and SijN_hyd from the mobility form, solved directly.  SijN_hyd
        is the stresslet of the strain alone (Fe = 0), as stepper_nogmres's
        mu EIN is."""
        n6 = APP.shape[0]
        M, R, K = self.mobility_form(APP, APQ, AQQ, rfu, rfe, rse)
        x = np.linalg.solve(K, np.concatenate([Fe, -EIN]))
        U = (M @ x)[:n6]
        x0 = np.linalg.solve(K, np.concatenate([np.zeros(n6), -EIN]))
        S = (x0 + R @ (M @ x0))[n6:]
        return U, S, K

    @staticmethod
    def block_jacobi(K, N):
        """Inverses of K's 11x11 diagonal blocks (particle g: T 3g.., R 3N+3g..,
        Q 6N+5g..), as a dense block-diagonal matrix in field-major order."""
        P = np.zeros_like(K)
        for g in range(N):
            idx = np.r_[3 * g:3 * g + 3, 3 * N + 3 * g:3 * N + 3 * g + 3, 6 * N + 5 * g:6 * N + 5 * g + 5]
            P[np.ix_(idx, idx)] = np.linalg.inv(K[np.ix_(idx, idx)])
        return P

    @staticmethod
    def gmres_iterations(K, b, P=None, tol=1e-13, max_iter=2000):
        """Iterations of full (unrestarted) right-preconditioned GMRES from zero
        to ||b - K x|| <= tol ||b|| (modified Gram-Schmidt, Givens), the count
        the GPU's restarted flexible GMRES is compared against."""
        n = len(b)
        P = np.eye(n) if P is None else P
        bn = np.linalg.norm(b)
        V = np.zeros((n, max_iter + 1))
        H = np.zeros((max_iter + 1, max_iter))
        V[:, 0] = b / bn
        g = np.zeros(max_iter + 1)
        g[0] = bn
        cs, sn = np.zeros(max_iter), np.zeros(max_iter)
        for k in range(max_iter):
            w = K @ (P @ V[:, k])
            for i in range(k + 1):
                H[i, k] = V[:, i] @ w
                w -= H[i, k] * V[:, i]
            H[k + 1, k] = np.linalg.norm(w)
            V[:, k + 1] = w / H[k + 1, k]
            for i in range(k):
                a, c = H[i, k], H[i + 1, k]
                H[i, k], H[i + 1, k] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
            r = np.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / r, H[k + 1, k] / r
            H[k, k], H[k + 1, k] = r, 0.0
            g[k + 1], g[k] = -sn[k] * g[k], cs[k] * g[k]
            if abs(g[k + 1]) <= tol * bn:
                return k + 1
        return max_iter


# ------------------------------------------------------------- reference
class Reference:
    """The flang-built reference (oracle/_ref/libref.so) through ref_driver.f90."""

    def __init__(self, path: str | None = None, lubdat_parent: str | None = None):
        path = path or os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = self.lib = C.CDLL(path, mode=os.RTLD_LAZY)
        L.ref_init()
        self.lubdat_parent = lubdat_parent
        L.ref_setup.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_double, C.c_double]
        L.ref_set_positions.argtypes = [C.c_int, _dp]
        L.ref_set_lattice_skew.argtypes = [C.c_double, _dp]
        L.ref_calc_lattice_inverse.argtypes = [C.c_double]
        L.ref_get_lattice.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip]
        L.ref_set_ewald.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int]
        L.ref_set_radii_test.argtypes = [C.c_double]
        L.ref_per_skew_correction.argtypes = [_dp, _dp]
        L.ref_period_correction.argtypes = [_dp, C.c_double]
        L.ref_rp_self.argtypes = [C.c_double, _dp, _dp, _dp]
        six = [_dp] * 6
        L.ref_rp_ij.argtypes = [_dp, C.c_double, C.c_double] + six
        L.ref_rp_regular.argtypes = [_dp, C.c_double, C.c_double] + six
        L.ref_rp_free_fts.argtypes = [_dp, C.c_double, C.c_double] + six
        L.ref_rp_yamakawa.argtypes = [_dp, C.c_double, C.c_double] + six
        L.ref_overlap_yamakawa.argtypes = [_dp, C.c_double, C.c_double] + six
        L.ref_per_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L.ref_lubmxcalc.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.ref_matrev.argtypes = [C.c_int, _dp]
        L.ref_mob_to_resist_fts.argtypes = [C.c_int] + [_dp] * 6
        L.ref_arrangement.argtypes = [C.c_int] + [_dp] * 4
        L.ref_ein.argtypes = [C.c_int, _dp]
        L.ref_nogmres.argtypes = [C.c_int] + [_dp] * 11
        L.ref_stepper_stokesian.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.ref_get_ntdem.argtypes = [_ip, _dp]
        L.ref_step_replay.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ref_set_brownian.argtypes = [C.c_int, C.c_double, C.c_int]
        L.ref_cholesky.argtypes = [C.c_int, _dp]
        L.ref_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ref_set_geometry.argtypes = [C.c_int, C.c_int]
        L.ref_wall_grpery_inv_fri.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ref_read_control.argtypes = [_ip, _dp]
        L.ref_nondimensionalise.argtypes = [_dp]
        L.ref_read_config.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp]
        L.ref_vmd_write.argtypes = [C.c_int, _dp, C.c_double, _dp, _dp]
        L.ref_yetamu_write.argtypes = [C.c_double, _dp, C.c_double, C.c_double, C.c_double, C.c_double]

    def setup(self, conf, radii, LB, gamma=1.0, linerr=1e-4, mu_f=1.0 / PAI, correction_method=2, nb=0,
              gravity=0.0, dt_dem0=1.0, radii_test=1.0):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        LB = np.ascontiguousarray(LB, dtype=np.float64)
        self.N = len(radii)
        self.lib.ref_setup(self.N, nb, _p(conf), _p(radii), _p(LB), gamma, linerr, mu_f, correction_method,
                           gravity, dt_dem0)
        self.lib.ref_set_radii_test(radii_test)

    def set_geometry(self, is_period=True, wall=False):
        """IsPeriod = simplePeriod, Wall_method (ref_setup leaves periodic, no wall)."""
        self.lib.ref_set_geometry(int(is_period), int(wall))

    def set_brownian(self, on, kbt=0.0, seed=1):
        """BROWN, kBT, and the intrinsic RNG's seed (every seed element = seed)."""
        self.lib.ref_set_brownian(int(on), float(kbt), int(seed))

    def cholesky(self, A):
        """CHOLESKY (matrices.f90:126-157): lower factor, strict upper zeroed."""
        A = np.array(A, dtype=np.float64, order="F")
        with _quiet_stdout(self.lib.ref_flush):
            self.lib.ref_cholesky(A.shape[0], _p(A))
        return A

    def set_lattice_skew(self, T):
        lat = _f((3, 3))
        self.lib.ref_set_lattice_skew(T, _p(lat))
        return lat

    def calc_lattice_inverse(self, ews_err):
        self.lib.ref_calc_lattice_inverse(ews_err)

    def get_lattice(self):
        lr, li = _f((3, 3)), _f((3, 3))
        s = [np.zeros(1) for _ in range(5)]
        nr, ni = C.c_int(), C.c_int()
        self.lib.ref_get_lattice(_p(lr), _p(li), *[_p(x) for x in s], C.byref(nr), C.byref(ni))
        return dict(LR=lr, LI=li, LV=s[0][0], sig=s[1][0], lam=s[2][0], logerr=s[3][0], phi=s[4][0],
                    NR=nr.value, NI=ni.value)

    def set_ewald(self, lam, phi, NR, NI):
        self.lib.ref_set_ewald(lam, phi, NR, NI)

    def per_skew_correction(self, R):
        R = np.array(R, dtype=np.float64)
        d = np.zeros(1)
        self.lib.ref_per_skew_correction(_p(R), _p(d))
        return R, d[0]

    def period_correction(self, R, T):
        R = np.array(R, dtype=np.float64)
        self.lib.ref_period_correction(_p(R), T)
        return R

    def rp_self(self, a):
        tt, rr, qq = _f((3, 3)), _f((3, 3)), _f((5, 5))
        self.lib.ref_rp_self(a, _p(tt), _p(rr), _p(qq))
        return tt, rr, qq

    def _six(self, fn, R, ai, aj):
        R = np.ascontiguousarray(R, dtype=np.float64)
        o = [_f((3, 3)), _f((3, 3)), _f((3, 3)), _f((5, 5)), _f((3, 5)), _f((3, 5))]
        fn(_p(R), ai, aj, *[_p(x) for x in o])
        return o

    def rp_ij(self, R, ai, aj):
        return self._six(self.lib.ref_rp_ij, R, ai, aj)

    def rp_regular(self, R, ai, aj):
        return self._six(self.lib.ref_rp_regular, R, ai, aj)

    def rp_free_fts(self, R, ai, aj):
        return self._six(self.lib.ref_rp_free_fts, R, ai, aj)

    def rp_yamakawa(self, R, ai, aj):
        return self._six(self.lib.ref_rp_yamakawa, R, ai, aj)

    def overlap_yamakawa_correction(self, R, ai, aj):
        return self._six(self.lib.ref_overlap_yamakawa, R, ai, aj)

    def per_grpery_inv_fri(self, conf, radii, ews_err):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_per_grpery_inv_fri, N, _p(conf), _p(radii), ews_err, _p(APP), _p(APQ), _p(AQQ))
        return APP, APQ, AQQ

    def grpery_inv_fri(self, conf, radii):
        """GRPERY_INV_FRI (GRPY_YZD.f90:14-264), free space."""
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        APP, APQ, AQQ = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_grpery_inv_fri, N, _p(conf), _p(radii), _p(APP), _p(APQ), _p(AQQ))
        return APP, APQ, AQQ

    def wall_grpery_inv_fri(self, conf, radii):
        """Wall_GRPERY_INV_FRI (Wall_interaction.f90:29-149) -> TPP(3N,3N)."""
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        T = _f((3 * N, 3 * N))
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_wall_grpery_inv_fri, N, _p(conf), _p(radii), _p(T))
        return T

    @contextlib.contextmanager
    def _in_lubdat_parent(self):
        old = os.getcwd()
        if self.lubdat_parent:
            os.chdir(self.lubdat_parent)
        try:
            yield
        finally:
            os.chdir(old)

    def lubmxcalc(self, conf):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        N = conf.shape[0]
        rfu, rfe, rse = _f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))
        with self._in_lubdat_parent(), _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_lubmxcalc, N, _p(conf), _p(rfu), _p(rfe), _p(rse))
        return rfu, rfe, rse

    def matrev(self, A):
        A = np.array(A, dtype=np.float64, order="F")
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_matrev, A.shape[0], _p(A))
        return A

    def mob_to_resist_fts(self, APP, APQ, AQQ):
        N = APP.shape[0] // 6
        o = [_f((6 * N, 6 * N)), _f((6 * N, 5 * N)), _f((5 * N, 5 * N))]
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_mob_to_resist_fts, N, _p(np.asfortranarray(APP)), _p(np.asfortranarray(APQ)),
                                           _p(np.asfortranarray(AQQ)), *[_p(x) for x in o])
        return o

    def arrangement(self, RPP, RPQ, RQQ):
        N = RPP.shape[0] // 6
        mu = _f((5 * N, 5 * N))
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_arrangement, N, _p(np.asfortranarray(RPP)), _p(np.asfortranarray(RPQ)),
                                     _p(np.asfortranarray(RQQ)), _p(mu))
        return mu

    def step_replay(self, conf, radii, dt, T, u_pos=None):
        """ref_driver.f90 ref_step_replay: STEPPER_Stokesian replayed through the
        reference's own routines.  Returns (conf, U_pos, yeta_mu, NT_DEM, U_par,
        SijN_hyd), the last two of the last sub-step."""
        conf = np.array(conf, dtype=np.float64, order="C")
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        u_pos = np.zeros(6 * N) if u_pos is None else np.array(u_pos, dtype=np.float64)
        yeta, upar, shyd = np.zeros(5), np.zeros(6 * N), np.zeros(5 * N)
        self.last_z, self.last_sb = np.zeros(3 * N), np.zeros(5 * N)
        with self._in_lubdat_parent(), _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_step_replay, N, _p(conf), _p(radii), dt, T, _p(yeta), _p(u_pos), _p(upar),
                          _p(shyd), _p(self.last_z), _p(self.last_sb))
        nt, dtd = C.c_int(), np.zeros(1)
        self.lib.ref_get_ntdem(C.byref(nt), _p(dtd))
        return conf, u_pos, yeta, nt.value, upar, shyd

    # ---- on-disk formats (initial_conf.f90), run in the directory holding the files
    @contextlib.contextmanager
    def _in_dir(self, d):
        old = os.getcwd()
        os.chdir(d)
        try:
            yield
        finally:
            os.chdir(old)

    def read_control(self, d):
        """INIT_READ_CONTROL on d/control_file.dat -> (ints(24), reals(48)) in rsd_control's order."""
        ii, rr = np.zeros(24, dtype=np.int32), np.zeros(48)
        with self._in_dir(d), _quiet_stdout(self.lib.ref_flush):
            self.lib.ref_read_control(ii.ctypes.data_as(_ip), _p(rr))
        return ii, rr

    def nondimensionalise(self):
        """dimentionless_calc/_translate + main.f90:148-155 on the last control read."""
        o = np.zeros(10)
        self.lib.ref_nondimensionalise(_p(o))
        return o

    def read_config(self, d, nmax=4096):
        """INITIAL_SIZE + INITIAL_CONF on d/initial_config.dat -> (conf, radii, LB, T, u_pos)."""
        n = C.c_int()
        lb, t = np.zeros(3), np.zeros(1)
        conf, radii, up = np.zeros((nmax, 3)), np.zeros(nmax), np.zeros(6 * nmax)
        with self._in_dir(d), _quiet_stdout(self.lib.ref_flush):
            self.lib.ref_read_config(nmax, C.byref(n), _p(lb), _p(t), _p(conf), _p(radii), _p(up))
        N = n.value
        return conf[:N].copy(), radii[:N].copy(), lb, float(t[0]), np.concatenate([up[:3 * N], np.zeros(3 * N)])

    def write_outputs(self, d, frames, yrows):
        """VMD_WRITE each (conf, T, radii, u_pos) frame and yetamu_WRITE each
        (T, yeta_mu, LB1, frequency, GAMMA, GAMMAangle) row into d/output.VTF and
        d/yeta_mu.txt; returns the two files' text."""
        with self._in_dir(d), _quiet_stdout(self.lib.ref_flush):
            self.lib.ref_io_open()
            for conf, T, radii, up in frames:
                c = np.ascontiguousarray(conf, dtype=np.float64)
                self.lib.ref_vmd_write(len(radii), _p(c), T, _p(np.ascontiguousarray(radii, dtype=np.float64)),
                                       _p(np.ascontiguousarray(up, dtype=np.float64)))
            for T, y, lb1, fr, g, ga in yrows:
                self.lib.ref_yetamu_write(T, _p(np.ascontiguousarray(y, dtype=np.float64)), lb1, fr, g, ga)
            self.lib.ref_io_close()
        with open(os.path.join(d, "output.VTF")) as f1, open(os.path.join(d, "yeta_mu.txt")) as f2:
            return f1.read(), f2.read()

    def ein(self, N):
        e = np.zeros(5 * N)
        self.lib.ref_ein(N, _p(e))
        return e

    def nogmres(self, APP, APQ, AQQ, rfu, rfe, rse, Fe, EIN):
        N = APP.shape[0] // 6
        u, s, rpp = np.zeros(6 * N), np.zeros(5 * N), _f((6 * N, 6 * N))
        arrs = [np.asfortranarray(x) for x in (APP, APQ, AQQ, rfu, rfe, rse)]
        Fe = np.ascontiguousarray(Fe, dtype=np.float64)
        EIN = np.ascontiguousarray(EIN, dtype=np.float64)
        with _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_nogmres, N, *[_p(x) for x in arrs], _p(Fe), _p(EIN), _p(u), _p(s), _p(rpp))
        return u, s, rpp

    def stepper_stokesian(self, conf, radii, dt, T, u_pos=None):
        conf = np.array(conf, dtype=np.float64, order="C")
        radii = np.ascontiguousarray(radii, dtype=np.float64)
        N = len(radii)
        u_pos = np.zeros(6 * N) if u_pos is None else np.array(u_pos, dtype=np.float64)
        yeta = np.zeros(5)
        with self._in_lubdat_parent(), _quiet_stdout(self.lib.ref_flush):
            _on_big_stack(self.lib.ref_stepper_stokesian, N, _p(conf), _p(radii), dt, T, _p(yeta), _p(u_pos))
        nt, dtd = C.c_int(), np.zeros(1)
        self.lib.ref_get_ntdem(C.byref(nt), _p(dtd))
        return conf, u_pos, yeta, nt.value


# ------------------------------------------------------------- full step
@dataclass
class StepConfig:
    """The reference's control settings that the periodic FTS step reads
    (control_file.dat order: initial_conf.f90:116-179; globals: mymodules.f90)."""

    LB: np.ndarray
    gamma: float = 1.0            # SYS_property::GAMMA (shear rate)
    linerr: float = 1e-4          # LATTICE_BASE::LINERR (EWS_ERR)
    mu_f: float = 1.0 / PAI       # forced to 1/pi by nondimensionalisation (mymodules.f90:816,846)
    correction_method: int = 2
    nb: int = 0
    gravity: float = 0.0
    dt_dem0: float = 1.0
    radii_test: float = 1.0
    u_bg: np.ndarray = field(default_factory=lambda: np.zeros(3))
    omega_bg: np.ndarray = field(default_factory=lambda: np.zeros(3))
    ei_bg: np.ndarray = field(default_factory=lambda: np.zeros(6))
    is_period: bool = True        # IsPeriod = simplePeriod (method module)
    wall: bool = False            # Wall_method
    kbt: float = 0.0              # BROWN with this kBT when > 0 (BROWN_proerty)


def _oracle_step(self: Oracle, cfg: StepConfig, conf, radii, dt, T, z=None):
    """stepper.f90:33-156 STEPPER_Stokesian on the periodic FTS path
    (IsPeriod, simplePeriod, FTS_method, uselub; no Brownian, walls, DEM
    stepper, GMRES, filaments or swimmers).  Returns (conf, U_pos, yeta_mu,
    NT_DEM, extras) where extras holds U_par and SijN_hyd."""
    conf = np.array(conf, dtype=np.float64)
    radii = np.asarray(radii, dtype=np.float64)
    N = len(radii)
    # INIT_u_bg (csd_post.f90:442-479)
    ei_bg = np.array(cfg.ei_bg, dtype=np.float64)
    ei_bg[4] = cfg.gamma * 0.5
    Eij = np.array([[ei_bg[0], ei_bg[3], ei_bg[4]], [ei_bg[3], ei_bg[1], ei_bg[5]], [ei_bg[4], ei_bg[5], ei_bg[2]]])
    omega_bg = np.array(cfg.omega_bg, dtype=np.float64)
    omega_bg[1] = cfg.gamma * 0.5
    EIN = np.tile(self.ei_transy2(Eij), N)                               # tensors.f90:905-919
    # pos_collision_judge / ppiclf_collision_timestep (csd_post.f90:383-436)
    if self.pos_collision(conf, radii, cfg.nb):
        nt_dem = 5 if dt < cfg.dt_dem0 else min(int(np.floor(dt / cfg.dt_dem0)), 100)
    else:
        nt_dem = 1
    dt_dem = dt / nt_dem
    lat = self.lattice(T, cfg.LB, cfg.gamma)                            # SET_LATTICE_SKEW
    if cfg.is_period:
        APP, APQ, AQQ, _ = self.per_grpery_inv_fri(conf, radii, lat, cfg.linerr, cfg.correction_method, cfg.mu_f)
    else:                                                                # GRPERY_INV_FRI (free space)
        APP, APQ, AQQ, _ = self.grpery_inv_fri(conf, radii, cfg.correction_method, cfg.mu_f)
        self.calc_lattice_inverse(lat, cfg.linerr, radii)
    if cfg.wall:                                                         # stepper.f90:67-70
        APP = APP.copy(order="F")
        APP[:3 * N, :3 * N] += self.wall_grpery_inv_fri(conf, radii, cfg.mu_f)
    # Brownian velocity (stepper.f90:72-86): CHOLESKY of APP's TT block
    # (matrices.f90:126-157), z given (GAUSS draws it in the reference), and
    # DMDX_CAL (stepper.f90:650-688): two grand mobilities at conf +- deta z / 2
    UB = np.zeros(6 * N)
    if cfg.kbt > 0.0:
        z = np.asarray(z, dtype=np.float64)
        A = np.linalg.cholesky(APP[:3 * N, :3 * N])
        VB = np.sqrt(2.0 * cfg.kbt) * (A @ z) / dt
        deta = 0.001
        dz = z.reshape(N, 3)
        mats = []
        for c in (conf + 0.5 * deta * dz, conf - 0.5 * deta * dz):
            if cfg.is_period:
                mats.append(self.per_grpery_inv_fri(c, radii, lat, cfg.linerr, cfg.correction_method, cfg.mu_f)[0])
            else:
                mats.append(self.grpery_inv_fri(c, radii, cfg.correction_method, cfg.mu_f)[0])
        DMDX = (mats[0][:3 * N, :3 * N] - mats[1][:3 * N, :3 * N]) @ z / deta
        UB[:3 * N] = VB + cfg.kbt * DMDX
    S_B = np.zeros(5 * N)
    Np = N - cfg.nb
    for _ in range(nt_dem):
        Fe = np.zeros(6 * N)                                             # source_body_F
        Fe[2:3 * (N - cfg.nb):3] = -cfg.gravity
        rfu, rfe, rse, _ = self.lubmxcalc(conf, radii, lat, cfg.mu_f, cfg.nb, cfg.is_period)
        U_par, S_hyd, RPP = self.stepper_nogmres(APP, APQ, AQQ, rfu, rfe, rse, Fe, EIN)
        S_B = self.last_RPQ.T @ UB
        # STEPPER_conf (stepper.f90:160-227) with the pre-step T for the LE wrap
        uo = self.calc_uo_bg(conf, cfg.u_bg, omega_bg, Eij)
        U_pos = U_par + uo
        conf[:N - cfg.nb] += dt_dem * U_pos[:3 * (N - cfg.nb)].reshape(-1, 3)
        if cfg.is_period:
            for i in range(Np):
                conf[i] = self.period_correction(conf[i], cfg.LB, T, cfg.gamma)
            uo = self.calc_uo_bg(conf, cfg.u_bg, omega_bg, Eij)
            U_pos = U_par + uo
    # yetamu_solve (csd_post.f90:313-367) with SijN_FP = SijN_B = 0
    yeta = np.zeros(5)
    yeta[0] = np.sum((S_hyd[3::5] + S_B[3::5]) / (-2.0 * EIN[3::5])) / lat.LV / cfg.mu_f
    yeta[3] = np.sum(S_B[3::5] / (-2.0 * EIN[3::5])) / lat.LV / cfg.mu_f
    yeta[1] = lat.phi
    F = np.zeros(6 * N)
    F[0:3 * N:3] = 6.0 * PAI * cfg.mu_f * cfg.radii_test                # averageU matrices.f90:230-253
    UU = APP @ F
    yeta[2] = 1.0 / (np.sum(UU[0:3 * N:3]) / N)
    return conf, U_pos, yeta, nt_dem, dict(U_par=U_par, S_hyd=S_hyd, EIN=EIN, S_B=S_B, UB=UB)


Oracle.stepper_stokesian = _oracle_step
