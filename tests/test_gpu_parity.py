"""GPU parity of the HIP path against the oracle (C restatement of the
reference), called through the C ABI.  Tolerances: 1e-12 relative for the
assembled matrices (same algebra, different summation order) and 1e-10
relative for solve outputs (velocities, stresslets, viscosity), the bar the
FP64 path is held to."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _config(n, phi, seed, bidisperse=False):
    from rsd_amd.configs import bidisperse_radii, random_suspension
    radii = bidisperse_radii(n, seed) if bidisperse else 1.0
    return random_suspension(n, phi, seed, radii=radii)


@pytest.mark.parametrize("n", [50, 200, 704, 2048])
def test_matrev_spd(rsd, n):
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, n))
    A = X @ X.T / n + np.eye(n)
    Au = np.triu(A) + np.tril(rng.normal(size=(n, n)), -1)  # garbage below: only the upper triangle is read
    inv, info = rsd.matrev(np.asfortranarray(Au))
    assert info == 0
    ref = np.linalg.inv(A)
    assert relerr(ref, inv) < 1e-12
    assert np.allclose(inv, inv.T, rtol=0, atol=0)


def test_matrev_not_spd_reports_info(rsd):
    A = np.eye(80)
    A[70, 70] = -1.0
    out, info = rsd.matrev(np.asfortranarray(A))
    assert info == 71
    # the failed pivot propagates as NaN into the result (no silent garbage)
    assert np.isnan(out).any()


@pytest.mark.parametrize("n,phi,T,bi", [(8, 0.15, 0.0, False), (8, 0.15, 0.7, True), (64, 0.2, 0.0, False),
                                         (64, 0.2, 3.3, False)])
def test_grand_mobility(rsd, oracle, n, phi, T, bi):
    conf, radii, L = _config(n, phi, 11 + n, bi)
    LB = np.full(3, L)
    lat = rsd.set_lattice_skew(T, LB, 1.0)
    A = rsd.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    olat = oracle.lattice(T, LB, 1.0)
    B = oracle.per_grpery_inv_fri(conf, radii, olat, 1e-4)
    assert (lat.nr, lat.ni) == (olat.NR, olat.NI)
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12


@pytest.mark.parametrize("corr", [2, 1])
def test_grand_mobility_overlap_correction(rsd, oracle, corr):
    """Overlapping pairs with the regularised (2) and the Yamakawa (1) overlap
    correction (per_GRPY_YZD.f90:98-116), unequal radii and one sphere deep
    inside another included."""
    conf, radii, L = _config(12, 0.2, 5, bidisperse=True)
    conf[1] = conf[0] + np.array([1.7, 0.3, -0.2])  # overlapping pairs (r < a_i + a_j)
    conf[2] = conf[3] + np.array([-1.95, 0.1, 0.0])
    k = int(np.flatnonzero(radii != radii[0])[-1])
    conf[k] = conf[0] + np.array([0.05, -0.1, 0.08])  # r < |a_i - a_j|: one sphere inside the other
    LB = np.full(3, L)
    lat = rsd.set_lattice_skew(0.25, LB, 1.0)
    A = rsd.per_grpery_inv_fri(conf, radii, lat, 1e-4, correction_method=corr)
    B = oracle.per_grpery_inv_fri(conf, radii, oracle.lattice(0.25, LB, 1.0), 1e-4, correction_method=corr)
    assert A[3] == B[3] >= 3
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12


def test_free_space_yamakawa_vs_oracle(rsd, oracle):
    """GRPERY_INV_FRI with correction_method = 1: overlapping pairs take
    YAMAKAWA_IJ's blocks (GRPY_YZD.f90:148-150)."""
    from rsd_amd.configs import bidisperse_radii, random_suspension
    conf, radii, L = random_suspension(40, 0.3, 3, radii=bidisperse_radii(40, 3))
    conf[3] = conf[2] + np.array([0.3, 1.6, 0.1])
    k = int(np.flatnonzero(radii != radii[6])[-1])
    conf[k] = conf[6] + np.array([0.02, 0.05, -0.03])  # one sphere inside the other
    A = rsd.grpery_inv_fri(conf, radii, correction_method=1)
    B = oracle.grpery_inv_fri(conf, radii, correction_method=1)
    assert A[3] == B[3] >= 2
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12


def test_step_with_yamakawa_correction_reports_lapack_info(rsd, oracle):
    """A step with correction_method = 1 and a pair at gap -5e-4.  The
    reference's Yamakawa correction keeps stresslet terms at contact (its
    QQ, GQ, HQ blocks do not meet the free RPY's there), and with them the
    11N grand mobility of any overlapping configuration tried is indefinite
    (smallest eigenvalue about -0.037 here).  MATREV's DPOTRF('U') then
    reports INFO > 0 and the reference carries on with DPOTRI's output; the
    step returns that same INFO, taken from LAPACK on the oracle's grand
    mobility (itself held to the reference's PER_GRPERY_INV_FRI in
    tests/test_yamakawa.py)."""
    import scipy.linalg as sl
    conf, radii, L = _config(64, 0.35, 91)
    conf = _squeeze_closest_pair(conf, L, 1.9995)
    LB = np.full(3, L)
    APP, APQ, AQQ, nov = oracle.per_grpery_inv_fri(conf, radii, oracle.lattice(0.0, LB, 1.0), 1e-4,
                                                   correction_method=1)
    G = np.block([[APP, APQ], [APQ.T, AQQ]])
    lapack_info = sl.lapack.dpotrf(G, lower=0)[1]
    assert nov == 1 and lapack_info > 0
    sus = rsd.Suspension(conf, radii, rsd.default_config(L, correction_method=1))
    y, info = sus.step(5e-4, 0.0)
    print(f"Yamakawa step: INFO {info}, LAPACK DPOTRF('U') INFO {lapack_info}", flush=True)
    assert info == lapack_info
    sus.close()


@pytest.mark.parametrize("n,phi,T,bi", [(16, 0.3, 0.0, False), (16, 0.3, 0.4, True), (64, 0.4, 1.2, False)])
def test_lubrication(rsd, oracle, n, phi, T, bi):
    conf, radii, L = _config(n, phi, 21 + n, bi)
    LB = np.full(3, L)
    lat = rsd.set_lattice_skew(T, LB, 1.0)
    A = rsd.lubmxcalc(conf, radii, lat)
    B = oracle.lubmxcalc(conf, radii, oracle.lattice(T, LB, 1.0))
    assert A[3] == B[3]
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12


def test_lubrication_contact_reflection(rsd, oracle):
    conf, radii, L = _config(10, 0.2, 3)
    conf[4] = conf[5] + np.array([1.9, 0.2, 0.1])
    LB = np.full(3, L)
    lat = rsd.set_lattice_skew(0.0, LB, 1.0)
    A = rsd.lubmxcalc(conf, radii, lat)
    B = oracle.lubmxcalc(conf, radii, oracle.lattice(0.0, LB, 1.0))
    assert A[3] == B[3] >= 1
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12


@pytest.mark.parametrize("n", [8, 32])
def test_stepper_nogmres(rsd, oracle, n):
    conf, radii, L = _config(n, 0.3, 31 + n)
    LB = np.full(3, L)
    olat = oracle.lattice(0.1, LB, 1.0)
    M = oracle.per_grpery_inv_fri(conf, radii, olat, 1e-4)[:3]
    Lb = oracle.lubmxcalc(conf, radii, olat)[:3]
    ein = np.tile(oracle.ei_transy2(np.array([[0, 0, .5], [0, 0, 0], [.5, 0, 0]])), n)
    fe = np.zeros(6 * n)
    fe[2:3 * n:3] = -0.3
    u_o, s_o, rpp_o = oracle.stepper_nogmres(*M, *Lb, fe, ein)
    u_g, s_g, rpp_g = rsd.stepper_nogmres(*M, *Lb, fe, ein)
    assert relerr(rpp_o, rpp_g) < 1e-12
    assert relerr(u_o, u_g) < 1e-10
    assert relerr(s_o, s_g) < 1e-10


@pytest.mark.parametrize("n,phi,steps,dt,bi", [(16, 0.3, 4, 0.02, False), (32, 0.35, 3, 0.01, True)])
def test_step_trajectory(rsd, oracle, n, phi, steps, dt, bi):
    from rsd_oracle import StepConfig
    conf, radii, L = _config(n, phi, 41 + n, bi)
    cfg_o = StepConfig(LB=np.full(3, L))
    cfg_g = rsd.default_config(L)
    sus = rsd.Suspension(conf, radii, cfg_g)
    c_o = conf.copy()
    T = 0.0
    for _ in range(steps):
        c_o, u_o, y_o, nt_o, ex = oracle.stepper_stokesian(cfg_o, c_o, radii, dt, T)
        y_g, info = sus.step(dt, T)
        c_g, u_g = sus.state()
        up_g, s_g, nt_g = sus.last()
        assert info == 0 and nt_g == nt_o
        assert relerr(c_o, c_g) < 1e-10
        assert relerr(u_o, u_g) < 1e-10
        assert relerr(ex["S_hyd"], s_g) < 1e-10
        assert abs(y_o[0] - y_g[0]) <= 1e-10 * abs(y_o[0])
        assert y_o[1] == pytest.approx(y_g[1], rel=1e-14)
        T += dt
    sus.close()


def test_step_repeatable(rsd):
    conf, radii, L = _config(24, 0.3, 77)
    out = []
    for _ in range(2):
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        sus.step(0.01, 0.0)
        out.append(np.concatenate([sus.state()[0].ravel(), sus.last()[1]]))
        sus.close()
    assert np.array_equal(out[0], out[1])


def _squeeze_closest_pair(conf, L, dist):
    """Move the closest pair that does not straddle the box (raw and
    minimum-image distances agree) together along its line of centres to
    separation `dist`.  Each particle moves by less than the packing's
    minimum gap, so no other pair comes into contact."""
    n = len(conf)
    best, bi, bj = np.inf, -1, -1
    for i in range(n):
        d = conf[i + 1:] - conf[i]
        raw = np.sqrt(np.sum(d * d, axis=1))
        dm = d - L * np.round(d / L)
        img = np.sqrt(np.sum(dm * dm, axis=1))
        ok = np.abs(raw - img) < 1e-12
        if np.any(ok):
            k = np.argmin(np.where(ok, raw, np.inf))
            if raw[k] < best:
                best, bi, bj = raw[k], i, i + 1 + k
    conf = conf.copy()
    c = 0.5 * (conf[bi] + conf[bj])
    u = (conf[bj] - conf[bi]) / best
    conf[bi], conf[bj] = c - 0.5 * dist * u, c + 0.5 * dist * u
    return conf


@pytest.mark.parametrize("mode,gravity", [(0, 0.0), (1, 0.0), (1, 0.3), (2, 0.0), (2, 0.3)])
def test_step_collision_substeps(rsd, oracle, mode, gravity):
    """A step with an overlapping pair: NT_DEM = 5 collision sub-steps, each
    with its own lubrication and solve (the path every bench step takes).
    mode 0 factors RPP at every sub-step; mode 1 (the default) solves
    sub-steps 2-5 by CG preconditioned with the first sub-step's inverse, one
    launch per CG phase; mode 2 is the same CG in one cooperative launch per
    solve.  All are held to 1e-10 against the oracle's direct solves."""
    from rsd_oracle import StepConfig
    conf, radii, L = _config(64, 0.35, 91)
    conf = _squeeze_closest_pair(conf, L, 1.9995)
    LB = np.full(3, L)
    APP, APQ, AQQ, _ = oracle.per_grpery_inv_fri(conf, radii, oracle.lattice(0.0, LB, 1.0), 1e-4)
    assert np.linalg.eigvalsh(np.block([[APP, APQ], [APQ.T, AQQ]])).min() > 0  # a well-posed step
    cfg_o = StepConfig(LB=LB, gravity=gravity)
    sus = rsd.Suspension(conf, radii, rsd.default_config(L, gravity=gravity))
    sus.set_substep_solver(mode)
    c_o, T, dt = conf.copy(), 0.0, 5e-4
    for _ in range(2):
        c_o, u_o, y_o, nt_o, ex = oracle.stepper_stokesian(cfg_o, c_o, radii, dt, T)
        y_g, info = sus.step(dt, T)
        c_g, u_g = sus.state()
        up_g, s_g, nt_g = sus.last()
        assert nt_o == 5 and nt_g == 5 and info == 0
        iters, fallbacks = sus.substep_stats()
        if mode in (1, 2):
            assert iters > 0 and fallbacks == 0
        else:
            assert iters == 0
        assert relerr(c_o, c_g) < 1e-10
        assert relerr(u_o, u_g) < 1e-10
        assert relerr(ex["S_hyd"], s_g) < 1e-10
        assert abs(y_o[0] - y_g[0]) <= 1e-10 * abs(y_o[0])
        T += dt
    sus.close()


def test_substep_cg_fallback_is_the_direct_solve(rsd):
    """With no CG iterations allowed every later sub-step falls back to the
    factorisation, which must reproduce the direct mode bit for bit."""
    conf, radii, L = _config(64, 0.35, 91)
    conf = _squeeze_closest_pair(conf, L, 1.9995)
    out = []
    for mode, max_iter in [(0, 60), (1, 0), (2, 0)]:
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        sus.set_substep_solver(mode, 1e-14, max_iter)
        sus.step(5e-4, 0.0)
        iters, fallbacks = sus.substep_stats()
        assert fallbacks == (4 if mode else 0)
        c, u = sus.state()
        out.append(np.concatenate([c.ravel(), u, sus.last()[1]]))
        sus.close()
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])


@pytest.mark.parametrize("route", [0, 1])
def test_resistance_routes(rsd, oracle, route):
    """The resistance blocks by the explicit inverse of G (route 0, the
    reference's MobToResist_FTS) and by the Schur complement of its stresslet
    block (route 1, the default), with direct sub-steps, against the oracle."""
    from rsd_oracle import StepConfig
    conf, radii, L = _config(64, 0.35, 91)
    conf = _squeeze_closest_pair(conf, L, 1.9995)
    cfg_o = StepConfig(LB=np.full(3, L))
    sus = rsd.Suspension(conf, radii, rsd.default_config(L))
    sus.set_resistance_route(route)
    sus.set_substep_solver(0)
    c_o, T, dt = conf.copy(), 0.0, 5e-4
    for _ in range(2):
        c_o, u_o, y_o, nt_o, ex = oracle.stepper_stokesian(cfg_o, c_o, radii, dt, T)
        y_g, info = sus.step(dt, T)
        c_g, u_g = sus.state()
        up_g, s_g, nt_g = sus.last()
        assert nt_o == nt_g == 5 and info == 0
        assert relerr(c_o, c_g) < 1e-10
        assert relerr(u_o, u_g) < 1e-10
        assert relerr(ex["S_hyd"], s_g) < 1e-10
        assert abs(y_o[0] - y_g[0]) <= 1e-10 * abs(y_o[0])
        T += dt
    sus.close()


def test_schur_route_redoes_an_indefinite_step(rsd):
    """A planted deep overlap makes the grand mobility indefinite.  The
    Schur route then redoes the step through the explicit inverse, so INFO
    (the order of G's first non-positive leading minor) and every output match
    the explicit route bit for bit (NaN where the failed pivot propagates)."""
    conf, radii, L = _config(16, 0.3, 3)
    conf[1] = conf[0] + np.array([1.0, 0.0, 0.0])
    out = []
    for route, mode in [(0, 0), (1, 1)]:
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        sus.set_resistance_route(route)
        sus.set_substep_solver(mode)
        y, info = sus.step(5e-4, 0.0)
        c, u = sus.state()
        out.append((info, np.concatenate([c.ravel(), u, sus.last()[1], y])))
        sus.close()
    assert out[0][0] > 0 and out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1], equal_nan=True)


def test_substep_cg_bench_workload(rsd):
    """N=512, phi=0.4 (the bench suspension) with its closest pair squeezed
    into contact: two NT_DEM=5 steps on the default path (Schur route, CG
    sub-steps), the one-launch CG and the reference's arithmetic (explicit
    inverse, factorisation at every sub-step), to 1e-10 (no oracle at this
    size: the paths share nothing after the grand mobility)."""
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(512, 0.4, 2000)
    conf = _squeeze_closest_pair(conf, L, 1.9995)
    res = []
    for route, mode in [(0, 0), (1, 1), (1, 2)]:
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        sus.set_resistance_route(route)
        sus.set_substep_solver(mode)
        T, rows = 0.0, []
        for _ in range(2):
            y, info = sus.step(5e-4, T)
            assert info == 0 and sus.last()[2] == 5
            if mode:
                iters, fallbacks = sus.substep_stats()
                print(f"route {route} mode {mode}: CG iterations {iters}, fallbacks {fallbacks}", flush=True)
                assert fallbacks == 0 and 0 < iters <= 4 * 60
            c, u = sus.state()
            rows.append((c, u, sus.last()[1], y[0]))
            T += 5e-4
        res.append(rows)
        sus.close()
    for (c0, u0, s0, y0), (c1, u1, s1, y1), (c2, u2, s2, y2) in zip(*res):
        assert relerr(c2, c1) < 1e-13 and relerr(u2, u1) < 1e-11   # the two CG forms: same iteration
        assert relerr(c0, c1) < 1e-12
        assert relerr(u0, u1) < 1e-10
        assert relerr(s0, s1) < 1e-10
        assert abs(y0 - y1) <= 1e-10 * abs(y0)
