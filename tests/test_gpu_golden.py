"""The HIP path against the reference's own outputs (golden.npz, written from
the flang build of src/*.f90 by tests/golden/make_golden.py), through the C ABI.

Covers the branches a random suspension does not reach: the Lees-Edwards wrap
of STEPPER_conf / PERIOD_CORRECTION (particles crossing both z faces and
x = L, with the pre-step T, across the strain reset), bidisperse lubrication
(s = r / r_ave, hydro_lub.f90:83-84) and boundary particles (Nb > 0).
Tolerances: 1e-12 relative for assembled matrices, 1e-10 for step outputs.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(GOLDEN, "golden.npz"))


@pytest.mark.parametrize("tag,nb", [("lub_", 0), ("lubbi_", 0), ("lubnb_", 2)])
def test_lubrication_golden(rsd, tag, nb):
    conf, radii = G[tag + "conf"], G[tag + "radii"]
    L, T = G[tag + "LT"]
    lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
    rfu, rfe, rse, nc = rsd.lubmxcalc(conf, radii, lat, nb=nb)
    print(f"lubmxcalc {tag} N={len(radii)} nb={nb}: contacts {nc}", flush=True)
    assert relerr(G[tag + "rfu"], rfu) < 1e-12
    assert relerr(G[tag + "rfe"], rfe) < 1e-12
    assert relerr(G[tag + "rse"], rse) < 1e-12


@pytest.mark.parametrize("tag", ["grmob8_", "grmobbi_"])
def test_grand_mobility_golden(rsd, tag):
    conf, radii = G[tag + "conf"], G[tag + "radii"]
    L, T = G[tag + "LT"]
    lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
    APP, APQ, AQQ, _ = rsd.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    assert relerr(G[tag + "APP"], APP) < 1e-12
    assert relerr(G[tag + "APQ"], APQ) < 1e-12
    assert relerr(G[tag + "AQQ"], AQQ) < 1e-12


def _wraps(prev, cur, L):
    d = np.abs(cur - prev)
    return int(np.sum(d[:, 2] > 0.5 * L)), int(np.sum((d[:, 0] > 0.5 * L) | (d[:, 1] > 0.5 * L)))


@pytest.mark.parametrize("tag,route,mode", [("trajle_", 1, 1), ("trajle_", 0, 0), ("trajnb_", 1, 1),
                                            ("traj_", 1, 1)])
def test_trajectory_golden(rsd, tag, route, mode):
    """Whole STEPPER_Stokesian steps against the reference: the Lees-Edwards
    trajectory must wrap particles on the GPU too (z and x), as it does in the
    reference."""
    if tag == "traj_":
        L, dt = G["traj_L_dt"]
        T0, nb = 0.0, 0
    else:
        L, dt, T0, nb = G[tag + "L_dt_T0_nb"]
        nb = int(nb)
    conf, radii = G[tag + "conf0"], G[tag + "radii"]
    sus = rsd.Suspension(conf, radii, rsd.default_config(L, nb=nb))
    sus.set_resistance_route(route)
    sus.set_substep_solver(mode)
    T, prev, zw, xw = T0, conf.copy(), 0, 0
    for k in range(len(G[tag + "ntdem"])):
        y, info = sus.step(dt, T)
        c, u = sus.state()
        up, s, nt = sus.last()
        dz, dx = _wraps(prev, c, L)
        zw, xw, prev = zw + dz, xw + dx, c
        print(f"{tag} step {k}: INFO={info} NT_DEM={nt} wraps z {zw} x {xw}", flush=True)
        assert info == 0 and nt == G[tag + "ntdem"][k]
        assert np.max(np.abs(c - G[tag + "conf"][k])) < 1e-10 * L
        assert relerr(G[tag + "upos"][k], u) < 1e-10
        if tag != "traj_":
            assert relerr(G[tag + "upar"][k], up) < 1e-10
            assert relerr(G[tag + "shyd"][k], s) < 1e-10
        assert relerr(G[tag + "yeta"][k][:3], y[:3]) < 1e-10
        T += dt
    sus.close()
    if tag == "trajle_":
        assert (zw, xw) == tuple(G["trajle_wraps"]) and zw >= 1


def test_free_space_grand_mobility_golden(rsd):
    """GRPERY_INV_FRI (GRPY_YZD.f90:14-264): free-space RPY, one regularised overlap."""
    conf, radii = G["free_conf"], G["free_radii"]
    APP, APQ, AQQ, nov = rsd.grpery_inv_fri(conf, radii)
    assert nov == 1
    assert relerr(G["free_APP"], APP) < 1e-12
    assert relerr(G["free_APQ"], APQ) < 1e-12
    assert relerr(G["free_AQQ"], AQQ) < 1e-12


def test_wall_mobility_golden(rsd):
    """Wall_GRPERY_INV_FRI (Wall_interaction.f90:29-149)."""
    T = rsd.wall_grpery_inv_fri(G["free_conf"], G["free_radii"])
    assert relerr(G["wall_TPP"], T) < 1e-12


@pytest.mark.parametrize("n,seed", [(40, 1), (130, 2)])
def test_free_space_and_wall_vs_oracle(rsd, oracle, n, seed):
    from rsd_amd.configs import bidisperse_radii, random_suspension
    conf, radii, L = random_suspension(n, 0.3, seed, radii=bidisperse_radii(n, seed))
    conf[3] = conf[2] + np.array([0.3, 2.2, 0.1])     # overlapping pair (radii <= 1.4 each)
    conf[:, 2] += 10.0 - 0.5 * L
    A = rsd.grpery_inv_fri(conf, radii)
    B = oracle.grpery_inv_fri(conf, radii)
    assert A[3] == B[3] >= 1
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12
    assert relerr(oracle.wall_grpery_inv_fri(conf, radii), rsd.wall_grpery_inv_fri(conf, radii)) < 1e-12


@pytest.mark.parametrize("route,mode", [(1, 1), (0, 0)])
def test_trajectory_free_space_with_wall_golden(rsd, route, mode):
    """rsd_ctx_set_geometry(is_period=0, wall_method=1): three steps against the
    reference's STEPPER_Stokesian with IsPeriod = simplePeriod = false and
    Wall_method on (U_pos, U_par, positions, yeta_mu(3))."""
    L, dt = G["trajfree_L_dt"]
    conf, radii = G["trajfree_conf0"], G["trajfree_radii"]
    sus = rsd.Suspension(conf, radii, rsd.default_config(L))
    sus.set_geometry(is_period=False, wall=True)
    sus.set_resistance_route(route)
    sus.set_substep_solver(mode)
    T = 0.0
    for k in range(3):
        y, info = sus.step(dt, T)
        c, u = sus.state()
        up, s, nt = sus.last()
        print(f"free+wall step {k}: INFO={info} NT_DEM={nt}", flush=True)
        assert info == 0 and nt == G["trajfree_ntdem"][k]
        assert np.max(np.abs(c - G["trajfree_conf"][k])) < 1e-10 * L
        assert relerr(G["trajfree_upos"][k], u) < 1e-10
        assert relerr(G["trajfree_upar"][k], up) < 1e-10
        assert abs(y[2] - G["trajfree_yeta"][k][2]) <= 1e-10 * abs(y[2])
        T += dt
    sus.close()


def test_cholesky_golden(rsd):
    L, info = rsd.cholesky(G["chol_in"])
    assert info == 0 and relerr(G["chol_out"], L) < 1e-12 and np.all(np.triu(L, 1) == 0.0)


def test_cholesky_reports_the_failed_minor(rsd):
    A = np.eye(70)
    A[66, 66] = -2.0
    L, info = rsd.cholesky(A)
    assert info == 67


@pytest.mark.parametrize("tag,is_period,route", [("trajbrown_", True, 1), ("trajbrown_", True, 0),
                                                 ("brownfree_", False, 1)])
def test_brownian_golden(rsd, tag, is_period, route):
    """The BROWN branch on the GPU with the reference generator's z: positions,
    velocities, SijN_hyd, the Brownian stresslet SijN_B and yeta_mu(1), (4)."""
    L, dt, T, kbt = G[tag + "L_dt_T0_kbt"]
    sus = rsd.Suspension(G[tag + "conf0"], G[tag + "radii"], rsd.default_config(L))
    sus.set_geometry(is_period=is_period)
    sus.set_resistance_route(route)
    sus.set_brownian(kbt)
    for k in range(len(G[tag + "ntdem"])):
        sus.set_brownian_z(G[tag + "z"][k])
        y, info = sus.step(dt, T)
        c, u = sus.state()
        up, s, nt = sus.last()
        z, ub, sb = sus.brownian()
        print(f"{tag} route {route} step {k}: INFO={info} NT_DEM={nt} yeta {y[0]:.6f} {y[3]:.6f}", flush=True)
        assert info == 0 and nt == G[tag + "ntdem"][k]
        assert np.array_equal(z, G[tag + "z"][k])
        assert np.max(np.abs(c - G[tag + "conf"][k])) < 1e-10 * L
        assert relerr(G[tag + "upos"][k], u) < 1e-10
        if is_period:
            assert relerr(G[tag + "sb"][k], sb) < 1e-10
            assert relerr(G[tag + "shyd"][k], s) < 1e-10
            assert relerr(G[tag + "yeta"][k][[0, 3]], y[[0, 3]]) < 1e-10
        assert abs(y[2] - G[tag + "yeta"][k][2]) <= 1e-10 * abs(y[2])
        T += dt
    sus.close()


def test_brownian_generator(rsd):
    """The device generator: standard normal moments over 3N = 3072 draws
    (each within four standard errors), repeatable for a seed, new numbers
    each step and for another seed."""
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(1024, 0.05, 3)
    zs = []
    for seed, steps in ((7, 2), (7, 1), (8, 1)):
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        sus.set_brownian(1e-3, seed)
        for _ in range(steps):
            sus.step(1e-3, 0.0)
            zs.append(sus.brownian()[0])
        sus.close()
    z = zs[0]
    n = len(z)
    assert abs(z.mean()) < 4.0 / np.sqrt(n) and abs(z.var() - 1.0) < 4.0 * np.sqrt(2.0 / n)
    assert abs(np.mean(z ** 4) - 3.0) < 4.0 * np.sqrt(96.0 / n) and abs(np.mean(z ** 3)) < 4.0 * np.sqrt(15.0 / n)
    assert np.array_equal(zs[0], zs[2])                       # same seed, same first step
    assert not np.array_equal(zs[0], zs[1]) and not np.array_equal(zs[0], zs[3])


def test_explicit_images_golden(rsd):
    """NR = NI = 4 given explicitly against the reference at the EWS_ERR whose
    rule gives 4 and 4."""
    conf, radii = G["img4_conf"], G["img4_radii"]
    L, T = G["img4_LT"]
    lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
    APP, APQ, AQQ, _ = rsd.per_grpery_inv_fri_images(conf, radii, lat, 1e-4, 4, 4, 0)
    assert (lat.nr, lat.ni) == (4, 4)
    assert relerr(G["img4_APP"], APP) < 1e-12
    assert relerr(G["img4_APQ"], APQ) < 1e-12
    assert relerr(G["img4_AQQ"], AQQ) < 1e-12


@pytest.mark.parametrize("nr,ni,k2", [(4, 6, 36), (2, 3, 5)])
def test_explicit_images_vs_oracle(rsd, oracle, nr, ni, k2):
    """The benchmark configurations' image set (4 real images, |n| <= 6) and a
    coarse spherical cutoff, grand mobility and one step, against the oracle."""
    from rsd_amd.configs import random_suspension
    from rsd_oracle import StepConfig
    conf, radii, L = random_suspension(48, 0.3, 17)
    # one regularised overlap (gap -5e-4) in a direction clear of the other
    # particles; at strain 0.05 the grand mobility stays positive definite
    # (at larger strains this unsheared configuration overlaps its
    # Lees-Edwards images and it does not, for the reference as well)
    u = np.array([-0.8753, 0.0402, -0.4819])
    conf[1] = conf[0] + 1.9995 * u / np.linalg.norm(u)
    T = 0.05
    lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
    A = rsd.per_grpery_inv_fri_images(conf, radii, lat, 1e-4, nr, ni, k2)
    oracle.set_images(nr, ni, k2)
    try:
        B = oracle.per_grpery_inv_fri(conf, radii, oracle.lattice(T, np.full(3, L), 1.0), 1e-4)
        c_o, u_o, y_o, nt_o, ex = oracle.stepper_stokesian(StepConfig(LB=np.full(3, L)), conf, radii, 5e-4, T)
    finally:
        oracle.set_images()
    assert np.linalg.eigvalsh(np.block([[B[0], B[1]], [B[1].T, B[2]]]))[0] > 0.0
    for x, y in zip(B[:3], A[:3]):
        assert relerr(x, y) < 1e-12
    sus = rsd.Suspension(conf, radii, rsd.default_config(L))
    sus.set_ewald_images(nr, ni, k2)
    y_g, info = sus.step(5e-4, T)
    c_g, u_g = sus.state()
    assert info == 0 and sus.last_nt() == nt_o
    assert relerr(u_o, u_g) < 1e-10 and relerr(ex["S_hyd"], sus.last()[1]) < 1e-10
    sus.close()
