"""Pins of the oracle from the mathematics (CPU).

Each check is a published closed form or an invariant the Ewald-summed RPY
mobility and the lubrication tables must satisfy, chosen so that a dropped
term, a wrong sign or index, or a transposed operand in the oracle fails one
of them.  The normalisation throughout is the reference's mu_f = 1/pi
(mymodules.f90:816,846), so the free-space Oseen prefactor 1/(8 pi mu r) is
1/(8 r) and the isolated-sphere self mobility 1/(6 pi mu a) is 1/(6 a).
"""
import os

import numpy as np
import pytest

from conftest import LUBDAT, relerr


def _rpy_free(R, ai, aj):
    """Rotne-Prager-Yamakawa mobilities of two non-overlapping spheres
    (r > ai + aj), unequal radii (Zuk, Wajnryb, Mizerski & Szymczak, JFM 741
    (2014) R5), in units mu = 1/pi:
      TT = 1/(8r) [(1 + (ai^2+aj^2)/(3r^2)) I + (1 - (ai^2+aj^2)/r^2) ee]
      RR = -1/(16 r^3) (I - 3 ee)
      TR(i,j) = 1/(8 r^2) eps_ijk e_k (up to the orientation of R)."""
    r = np.linalg.norm(R)
    e = R / r
    s = ai * ai + aj * aj
    ee = np.outer(e, e)
    I = np.eye(3)
    TT = (1.0 / (8 * r)) * ((1 + s / (3 * r * r)) * I + (1 - s / (r * r)) * ee)
    RR = -(1.0 / (16 * r ** 3)) * (I - 3 * ee)
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1
    TR = (1.0 / (8 * r * r)) * np.einsum("ijk,k->ij", eps, e)
    return TT, RR, TR


@pytest.mark.parametrize("ai,aj", [(1.0, 1.0), (1.0, 1.4), (1.4, 0.7)])
def test_free_pair_is_rpy(oracle, ai, aj):
    rng = np.random.default_rng(3)
    for _ in range(10):
        v = rng.normal(size=3)
        R = (ai + aj) * rng.uniform(1.01, 4.0) * v / np.linalg.norm(v)
        TT, RR, TR, *_ = oracle.rp_free_fts(R, ai, aj)
        tt, rr, tr = _rpy_free(R, ai, aj)
        assert relerr(tt, TT) < 1e-13
        assert relerr(rr, RR) < 1e-13
        # the sign of TR depends on whether R points i->j or j->i; the magnitude and structure do not
        assert min(relerr(tr, TR), relerr(-tr, TR)) < 1e-13


@pytest.mark.parametrize("L", [20.0, 40.0, 80.0])
def test_hasimoto_self_mobility(oracle, L):
    """One sphere in a simple-cubic periodic lattice (Hasimoto, JFM 5 (1959) 317):
    translational mobility (1/6a)(1 - 2.8372974794806 a/L + (4 pi/3)(a/L)^3 + O((a/L)^6)),
    rotational mobility (1/8a^3)(1 - (4 pi/3)(a/L)^3 + O((a/L)^6))."""
    a = 1.0
    lat = oracle.lattice(0.0, np.array([L, L, L]), 0.0)
    oracle.calc_lattice_inverse(lat, 1e-10, np.array([a]))
    APP, APQ, AQQ, _ = oracle.per_grpery_inv_fri(np.zeros((1, 3)), np.array([a]), lat, 1e-10)
    x = a / L
    tol = 30 * x ** 6
    mt = (1 / (6 * a)) * (1 - 2.8372974794806 * x + 4 * np.pi / 3 * x ** 3)
    mr = (1 / (8 * a ** 3)) * (1 - 4 * np.pi / 3 * x ** 3)
    assert np.allclose(np.diag(APP)[:3], mt, rtol=tol, atol=0)
    assert np.allclose(np.diag(APP)[3:], mr, rtol=tol, atol=0)
    # cubic symmetry: isotropic blocks, no translation-rotation coupling
    assert np.max(np.abs(APP - np.diag(np.diag(APP)))) < 1e-12
    assert np.max(np.abs(APQ)) < 1e-12


def test_ewald_splitting_independence(oracle):
    """The Ewald sum does not depend on the splitting parameter lambda once
    both the real-space and reciprocal-space cubes have converged."""
    rng = np.random.default_rng(5)
    LB = np.array([9.0, 10.0, 11.0])
    base = oracle.lattice(0.43, LB, 1.0)
    oracle.calc_lattice_inverse(base, 1e-4, np.array([1.0, 1.3]))
    results = []
    Rs = [rng.uniform(-4, 4, 3) for _ in range(4)]
    for scale in (0.7, 1.0, 1.5):
        lat = oracle.lattice(0.43, LB, 1.0)
        oracle.calc_lattice_inverse(lat, 1e-4, np.array([1.0, 1.3]))
        lat.lam = base.lam * scale
        lat.NR, lat.NI = 6, 8
        results.append([oracle.rp_ij(lat, R, 1.0, 1.3) for R in Rs])
    for other in results[1:]:
        for blocks0, blocks in zip(results[0], other):
            for b0, b in zip(blocks0, blocks):
                assert np.max(np.abs(b - b0)) < 1e-11 * max(1.0, np.max(np.abs(b0)))


def test_grand_mobility_symmetric_positive(oracle):
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(10, 0.2, seed=4)
    lat = oracle.lattice(0.3, np.full(3, L), 1.0)
    APP, APQ, AQQ, nov = oracle.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    M = np.block([[APP, APQ], [APQ.T, AQQ]])
    assert nov == 0
    assert np.max(np.abs(APP - APP.T)) < 1e-14 * np.max(np.abs(APP))
    assert np.max(np.abs(AQQ - AQQ.T)) < 1e-14 * np.max(np.abs(AQQ))
    assert np.linalg.eigvalsh(M).min() > 0


def test_matrev_is_inverse(oracle):
    rng = np.random.default_rng(9)
    X = rng.normal(size=(60, 60))
    A = X @ X.T / 60 + 0.5 * np.eye(60)
    Ai, info = oracle.matrev(A)
    assert info == 0
    assert np.max(np.abs(A @ Ai - np.eye(60))) < 1e-12
    assert np.array_equal(Ai, Ai.T)


def _tables():
    cols = {}
    for name in ("r2babc.dat", "r2bgh.dat", "r2bm.dat"):
        path = os.path.join(LUBDAT, name)
        with open(path) as f:
            header = f.readline().split()
        data = np.loadtxt(path, skiprows=1)
        for k, h in enumerate(header[1:], start=1):
            cols[h] = (data[:, 0], data[:, k])
    return cols


# oracle lubcalc order (hydro_lub.f90:310-525): the table column each entry comes from
LUB_ORDER = ["X11a", "X12a", "Y11a", "Y12a", "Y11b", "Y12b", "X11c", "X12c", "Y11c", "Y12c",
             "X11g", "X12g", "Y11g", "Y12g", "Y11h", "Y12h", "X11M+X12M", "Y11M+Y12M", "Z11M+Z12M"]


def test_lubcalc_hits_table_nodes(oracle):
    """At a tabulated separation the interpolation returns the tabulated value,
    column by column as the table headers name them (mobility sums halved)."""
    cols = _tables()
    for k, name in enumerate(LUB_ORDER):
        s, y = cols[name]
        for i in range(1, len(s) - 1):
            v = oracle.lubcalc(s[i])
            want = y[i] / 2 if name.endswith("M") else y[i]
            assert v[k] == pytest.approx(want, rel=1e-15, abs=1e-300), (name, s[i])
    assert np.all(oracle.lubcalc(4.5) == 0.0)  # beyond the tables: no lubrication


def test_lubcalc_near_contact_continuity(oracle):
    """The near-contact asymptotic forms (Jeffrey & Onishi 1984; Jeffrey 1992)
    used for s < 2.1 meet the tables at s = 2.1 to the tables' own accuracy;
    at contact X11A diverges as 1/(4 xi) with X12A -> -X11A."""
    below = oracle.lubcalc(2.1 - 1e-12)
    at = oracle.lubcalc(2.1)
    big = np.abs(at) > 0.05
    assert np.max(np.abs(below[big] - at[big]) / np.abs(at[big])) < 0.02
    xi = 1e-6
    v = oracle.lubcalc(2.0 + xi)
    assert v[0] * xi == pytest.approx(0.25, rel=1e-4)
    assert v[1] == pytest.approx(-v[0], rel=1e-6)
