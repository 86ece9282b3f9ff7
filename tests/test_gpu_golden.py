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
