"""The oracle against the reference's own outputs (CPU).

``golden.npz`` holds inputs and outputs of the reference Fortran (written by
``tests/golden/make_golden.py`` from the flang build of src/*.f90).  These
tests need neither the reference build nor a GPU.  When ``oracle/_ref`` is
present the live-reference tests below widen the same checks to fresh seeds.

Tolerance: 1e-12 relative to the largest entry of each output for single
evaluations; 1e-10 for the three-step trajectory (the FP64 parity bar of
the hot path).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, relerr

G = np.load(os.path.join(GOLDEN, "golden.npz"))
TOL = 1e-12


def _lat(oracle, T, LB, gamma, radii=None, ews_err=1e-4):
    lat = oracle.lattice(T, np.asarray(LB, dtype=np.float64), gamma)
    if radii is not None:
        oracle.calc_lattice_inverse(lat, ews_err, np.asarray(radii))
    return lat


def _unpack(v):
    """TT RR TR (3x3) QQ (5x5) GQ HQ (3x5), column-major, as make_golden packs them."""
    sizes = [(3, 3), (3, 3), (3, 3), (5, 5), (3, 5), (3, 5)]
    out, k = [], 0
    for s in sizes:
        n = s[0] * s[1]
        out.append(v[k:k + n].reshape(s, order="F"))
        k += n
    return out


def test_lattice(oracle):
    for row in G["lattice_rows"]:
        T, LB, gamma = row[0], row[1:4], row[4]
        lat = _lat(oracle, T, LB, gamma, radii=[1.0, 1.4])
        assert np.allclose(lat.lr.ravel(order="F"), row[5:14], rtol=0, atol=1e-14)
        assert np.allclose(lat.li.ravel(order="F"), row[14:23], rtol=0, atol=1e-14)
        LV, sig, lam, logerr, phi, NR, NI = row[23:30]
        assert abs(lat.LV - LV) <= 1e-12 * LV
        assert abs(lat.lam - lam) <= 1e-14 * lam and abs(lat.sig - sig) <= 1e-14 * sig
        assert abs(lat.logerr - logerr) <= 1e-14 * abs(logerr) and abs(lat.phi - phi) <= 1e-14
        assert (lat.NR, lat.NI) == (int(NR), int(NI))


def test_skew_correction(oracle):
    T, L1, L2, L3, gamma = G["skew_T_LB"]
    lat = _lat(oracle, T, [L1, L2, L3], gamma)
    for R, Rout, d in zip(G["skew_in"], G["skew_out"], G["skew_dmin"]):
        r, dm = oracle.per_skew_correction(lat, R)
        assert np.allclose(r, Rout, rtol=0, atol=1e-13)
        assert abs(dm - d) <= 1e-13


@pytest.mark.parametrize("which", ["ij", "regular", "free"])
def test_pair_tensors(oracle, which):
    T, L1, L2, L3, gamma = G["pair_lattice"]
    lat = _lat(oracle, T, [L1, L2, L3], gamma, radii=[1.0, 1.4])
    fn = {"ij": lambda R, a, b: oracle.rp_ij(lat, R, a, b), "regular": oracle.rp_regular,
          "free": oracle.rp_free_fts}[which]
    for R, (ai, aj), want in zip(G["pair_R"], G["pair_a"], G["pair_" + which]):
        got = fn(R, ai, aj)
        for g, w in zip(got, _unpack(want)):
            assert np.max(np.abs(g - w)) <= TOL * max(1.0, np.max(np.abs(w))), (which, R, ai, aj)


def test_self_tensors(oracle):
    T, L1, L2, L3, gamma = G["pair_lattice"]
    lat = _lat(oracle, T, [L1, L2, L3], gamma, radii=[1.0, 1.4])
    for a, want in zip((1.0, 1.4), G["pair_self"]):
        tt, rr, qq = oracle.rp_self(lat, a)
        assert relerr(want[:9], tt.ravel(order="F")) < TOL
        assert relerr(want[9:18], rr.ravel(order="F")) < TOL
        assert relerr(want[18:], qq.ravel(order="F")) < TOL


@pytest.mark.parametrize("tag", ["grmob8_", "grmobbi_"])
def test_grand_mobility(oracle, tag):
    conf, radii = G[tag + "conf"], G[tag + "radii"]
    L, T = G[tag + "LT"]
    lat = _lat(oracle, T, [L, L, L], 1.0)
    APP, APQ, AQQ, nov = oracle.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    assert relerr(G[tag + "APP"], APP) < TOL
    assert relerr(G[tag + "APQ"], APQ) < TOL
    assert relerr(G[tag + "AQQ"], AQQ) < TOL
    if tag == "grmob8_":
        assert nov == 2  # the two planted overlaps, nothing else


def test_lubrication(oracle):
    conf, radii = G["lub_conf"], G["lub_radii"]
    L, T = G["lub_LT"]
    lat = _lat(oracle, T, [L, L, L], 1.0)
    rfu, rfe, rse, _ = oracle.lubmxcalc(conf, radii, lat)
    assert relerr(G["lub_rfu"], rfu) < TOL
    assert relerr(G["lub_rfe"], rfe) < TOL
    assert relerr(G["lub_rse"], rse) < TOL


def test_matrev(oracle):
    inv, info = oracle.matrev(G["matrev_spd_in"])
    assert info == 0 and relerr(G["matrev_spd_out"], inv) < TOL
    # indefinite: DPOTRF stops at the 18th minor; the reference prints and goes on (matrices.f90:19-24).
    # What DPOTRI then returns depends on how far the LAPACK build's blocked factorisation got, so only
    # INFO is part of the contract (the HIP path returns INFO and fills the result with NaN).
    inv, info = oracle.matrev(G["matrev_ind_in"])
    assert info == 18
    assert np.all(np.isfinite(G["matrev_ind_out"]))


def test_mob_to_resist_and_arrangement(oracle):
    RPP, RPQ, RQQ, info = oracle.mob_to_resist_fts(G["lub_APP"], G["lub_APQ"], G["lub_AQQ"])
    assert info == 0
    assert relerr(G["m2r_RPP"], RPP) < 1e-11
    assert relerr(G["m2r_RPQ"], RPQ) < 1e-11
    assert relerr(G["m2r_RQQ"], RQQ) < 1e-11
    mu = oracle.arrangement(G["m2r_RPP"] + G["lub_rfu"], G["m2r_RPQ"] + G["lub_rfe"], G["m2r_RQQ"] + G["lub_rse"])
    assert relerr(G["m2r_mu"], mu) < 1e-11


def test_nogmres(oracle):
    U, S, RPP = oracle.stepper_nogmres(G["lub_APP"], G["lub_APQ"], G["lub_AQQ"], G["lub_rfu"], G["lub_rfe"],
                                       G["lub_rse"], G["nogm_fe"], G["nogm_ein"])
    assert relerr(G["nogm_RPP"], RPP) < 1e-11
    assert relerr(G["nogm_U"], U) < 1e-10
    assert relerr(G["nogm_S"], S) < 1e-10


def test_mobility_form_matches_nogmres(oracle):
    """The iterative route's single system K x = [Fe; -EIN] (the mobility form
    hydro_SD_FTS solves, stepper.f90:476-546) gives STEPPER_NOGmres's U_par and
    SijN_hyd (the golden nogm_* of the reference); dropping K's identity or
    negating its Q rows, the two plausible slips, does not."""
    args = [G[k] for k in ("lub_APP", "lub_APQ", "lub_AQQ", "lub_rfu", "lub_rfe", "lub_rse")]
    U, S, K = oracle.stepper_mobility_form(*args, G["nogm_fe"], G["nogm_ein"])
    assert relerr(G["nogm_U"], U) < 1e-10
    assert relerr(G["nogm_S"], S) < 1e-10
    n6 = G["lub_APP"].shape[0]
    M, R, _ = oracle.mobility_form(*args)
    b = np.concatenate([G["nogm_fe"], -G["nogm_ein"]])
    for bad in (K - np.pad(np.eye(n6), ((0, K.shape[0] - n6), (0, K.shape[0] - n6))),
                np.vstack([K[:n6], -K[n6:]])):
        x = np.linalg.solve(bad, b)
        assert relerr(G["nogm_U"], (M @ x)[:n6]) > 1e-3


def test_block_jacobi_reduces_gmres_iterations(oracle):
    """K's 11x11 block-Jacobi preconditioner cuts unpreconditioned GMRES's
    iteration count on the golden lubricated configuration."""
    args = [G[k] for k in ("lub_APP", "lub_APQ", "lub_AQQ", "lub_rfu", "lub_rfe", "lub_rse")]
    _, _, K = oracle.mobility_form(*args)
    b = np.concatenate([G["nogm_fe"], -G["nogm_ein"]])
    N = G["lub_APP"].shape[0] // 6
    plain = oracle.gmres_iterations(K, b)
    bj = oracle.gmres_iterations(K, b, oracle.block_jacobi(K, N))
    assert bj < plain <= K.shape[0]


def test_trajectory(oracle):
    from rsd_oracle import StepConfig
    L, dt = G["traj_L_dt"]
    conf, radii = G["traj_conf0"].copy(), G["traj_radii"]
    cfg = StepConfig(LB=np.full(3, L))
    T = 0.0
    for k in range(3):
        conf, upos, yeta, nt, _ = oracle.stepper_stokesian(cfg, conf, radii, dt, T)
        T += dt
        assert nt == G["traj_ntdem"][k] == 5
        assert np.max(np.abs(conf - G["traj_conf"][k])) < 1e-12 * L
        assert relerr(G["traj_upos"][k], upos) < 1e-10
        assert relerr(G["traj_yeta"][k][:3], yeta[:3]) < 1e-10


# ----------------------------------------------------------- live reference
def test_live_reference_pairs(oracle, reference):
    rng = np.random.default_rng(7)
    LB = np.array([9.0, 10.0, 11.0])
    radii = np.array([1.0, 1.3])
    reference.setup(np.zeros((2, 3)), radii, LB, gamma=1.0, linerr=1e-4)
    reference.set_lattice_skew(0.81)
    reference.calc_lattice_inverse(1e-4)
    lat = _lat(oracle, 0.81, LB, 1.0, radii=radii)
    for _ in range(20):
        R = rng.uniform(-5, 5, 3)
        ai, aj = rng.choice(radii, 2)
        for got, want in zip(oracle.rp_ij(lat, R, ai, aj), reference.rp_ij(R, ai, aj)):
            assert np.max(np.abs(got - want)) <= TOL * max(1.0, np.max(np.abs(want)))


def test_live_reference_step(oracle, reference):
    from rsd_oracle import StepConfig
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(12, 0.3, seed=99)
    reference.setup(conf, radii, np.full(3, L), gamma=1.0, linerr=1e-4)
    c_ref, u_ref, y_ref, nt_ref = reference.stepper_stokesian(conf, radii, 1e-3, 0.25)
    cfg = StepConfig(LB=np.full(3, L))
    c, u, y, nt, _ = oracle.stepper_stokesian(cfg, conf, radii, 1e-3, 0.25)
    assert nt == nt_ref
    assert np.max(np.abs(c - c_ref)) < 1e-12 * L
    assert relerr(u_ref, u) < 1e-10
    assert relerr(y_ref[:3], y[:3]) < 1e-10


# ---------------------------------------------- Lees-Edwards wrap, Nb, bidisperse
def test_lees_edwards_wrap(oracle):
    """PERIOD_CORRECTION (csd_post.f90:65-102): z crossings shift x by -/+ Lz gamma T
    (the unreduced T, then MOD Lx), x and y wrap once; rows as make_golden writes them."""
    rows = G["lewrap_rows"]
    assert len(rows) == 132
    moved = 0
    for row in rows:
        T, gamma, LB, R_in, R_out = row[0], row[1], row[2:5], row[5:8], row[8:11]
        got = oracle.period_correction(R_in, LB, T, gamma)
        assert np.max(np.abs(got - R_out)) <= 1e-13 * np.max(LB), (T, gamma, R_in)
        moved += int(np.any(R_out != R_in))
    assert moved == 120  # every row but the in-box one of each (LB, T) block wraps


def test_lees_edwards_wrap_mistakes_fail(oracle):
    """The golden rows tell a wrong sign of the Lz gamma T shift, or the time
    one step of 1e-3 later (the post-step T in place of the pre-step one,
    stepper.f90:213), from the reference's rule.  (Reducing T by the lattice's
    strain-reset period changes nothing: it shifts x by whole multiples of Lx.)"""
    rows = G["lewrap_rows"]
    bad_sign = bad_post = 0
    for row in rows:
        T, gamma, LB, R_in, R_out = row[0], row[1], row[2:5], row[5:8], row[8:11]
        if not (R_in[2] > LB[2] or R_in[2] < 0.0):
            continue
        got = oracle.period_correction(R_in, LB, -T, gamma)           # sign of the shift flipped
        bad_sign += int(np.max(np.abs(got - R_out)) > 1e-9)
        got = oracle.period_correction(R_in, LB, T + 1e-3, gamma)     # post-step T
        bad_post += int(np.max(np.abs(got - R_out)) > 1e-9)
    assert bad_sign == 60 and bad_post == 72  # 72 z crossings, 12 of them at T = 0


@pytest.mark.parametrize("tag,nb", [("lubbi_", 0), ("lubnb_", 2)])
def test_lubrication_bidisperse_and_boundary(oracle, tag, nb):
    conf, radii = G[tag + "conf"], G[tag + "radii"]
    L, T = G[tag + "LT"]
    lat = _lat(oracle, T, [L, L, L], 1.0)
    rfu, rfe, rse, nc = oracle.lubmxcalc(conf, radii, lat, nb=nb)
    assert relerr(G[tag + "rfu"], rfu) < TOL
    assert relerr(G[tag + "rfe"], rfe) < TOL
    assert relerr(G[tag + "rse"], rse) < TOL
    if tag == "lubbi_":
        assert nc == 1  # the planted s <= 2 pair is reflected


def test_lubrication_bidisperse_rule_matters(oracle):
    """Plant the plausible mistake (s = r / min(a_i, a_j) for the pair's mean
    radius) in the golden bidisperse case: the golden output rejects it."""
    conf, radii = G["lubbi_conf"], G["lubbi_radii"]
    L, T = G["lubbi_LT"]
    lat = _lat(oracle, T, [L, L, L], 1.0)
    rfu_mono = oracle.lubmxcalc(conf, np.full_like(radii, 1.0), lat)[0]
    assert relerr(G["lubbi_rfu"], rfu_mono) > 1e-3


@pytest.mark.parametrize("tag", ["trajle_", "trajnb_"])
def test_trajectory_wraps_and_boundary(oracle, tag):
    """Twelve bidisperse steps crossing both z faces and x = L (Lees-Edwards
    wraps with the pre-step T) and the strain reset at T = 1; three steps with
    Nb = 2 boundary particles.  Held to 1e-10 against the reference."""
    from rsd_oracle import StepConfig
    L, dt, T0, nb = G[tag + "L_dt_T0_nb"]
    nb = int(nb)
    conf, radii = G[tag + "conf0"].copy(), G[tag + "radii"]
    cfg = StepConfig(LB=np.full(3, L), nb=nb)
    T = T0
    for k in range(len(G[tag + "ntdem"])):
        conf, upos, yeta, nt, ex = oracle.stepper_stokesian(cfg, conf, radii, dt, T)
        T += dt
        assert nt == G[tag + "ntdem"][k]
        assert np.max(np.abs(conf - G[tag + "conf"][k])) < 1e-10 * L, k
        assert relerr(G[tag + "upos"][k], upos) < 1e-10, k
        assert relerr(G[tag + "upar"][k], ex["U_par"]) < 1e-10, k
        assert relerr(G[tag + "shyd"][k], ex["S_hyd"]) < 1e-10, k
        assert relerr(G[tag + "yeta"][k][:3], yeta[:3]) < 1e-10, k
    if tag == "trajle_":
        assert tuple(G["trajle_wraps"]) >= (1, 1)


# ------------------------------------------------------ free space and walls
def test_free_space_grand_mobility(oracle):
    """GRPERY_INV_FRI (GRPY_YZD.f90:14-264) against the reference, with one
    overlapping pair given the regularised RPY."""
    conf, radii = G["free_conf"], G["free_radii"]
    APP, APQ, AQQ, nov = oracle.grpery_inv_fri(conf, radii)
    assert nov == 1
    assert relerr(G["free_APP"], APP) < TOL
    assert relerr(G["free_APQ"], APQ) < TOL
    assert relerr(G["free_AQQ"], AQQ) < TOL


def test_wall_mobility(oracle):
    conf, radii = G["free_conf"], G["free_radii"]
    T = oracle.wall_grpery_inv_fri(conf, radii)
    assert relerr(G["wall_TPP"], T) < TOL
    below = conf[:, 2] <= 10.0
    assert 0 < below.sum() < len(radii)
    blocks = np.abs(G["wall_TPP"]).reshape(len(radii), 3, len(radii), 3).sum(axis=(1, 3))
    # a block is set exactly when one of its two particles is within z <= 10 (Wall_interaction.f90:50,89)
    want = below[:, None] | below[None, :]
    assert np.array_equal(blocks > 0, want)


def test_trajectory_free_space_with_wall(oracle):
    """Three steps with IsPeriod = simplePeriod = false and Wall_method on:
    free-space grand mobility plus the wall images' TT term, lubrication on raw
    separations, no Lees-Edwards wrap.  The reference forms no stresslet on
    this path (stepper.f90:400 needs IsPeriod), so U and the mobility-only
    yeta_mu(3) are compared."""
    from rsd_oracle import StepConfig
    L, dt = G["trajfree_L_dt"]
    conf, radii = G["trajfree_conf0"].copy(), G["trajfree_radii"]
    cfg = StepConfig(LB=np.full(3, L), is_period=False, wall=True)
    T = 0.0
    for k in range(3):
        conf, upos, yeta, nt, ex = oracle.stepper_stokesian(cfg, conf, radii, dt, T)
        T += dt
        assert nt == G["trajfree_ntdem"][k]
        assert np.max(np.abs(conf - G["trajfree_conf"][k])) < 1e-10 * L
        assert relerr(G["trajfree_upos"][k], upos) < 1e-10
        assert relerr(G["trajfree_upar"][k], ex["U_par"]) < 1e-10
        assert abs(yeta[2] - G["trajfree_yeta"][k][2]) <= 1e-10 * abs(yeta[2])


# ------------------------------------------------------------------ Brownian
def test_cholesky(oracle):
    """CHOLESKY (matrices.f90:126-157): lower factor, strict upper zeroed."""
    L = np.linalg.cholesky(G["chol_in"])
    assert relerr(G["chol_out"], L) < TOL
    assert np.all(np.triu(G["chol_out"], 1) == 0.0)


@pytest.mark.parametrize("tag,is_period", [("trajbrown_", True), ("brownfree_", False)])
def test_brownian_steps(oracle, tag, is_period):
    """BROWN (stepper.f90:72-86): CHOLESKY of APP's TT block, VB, the DMDX_CAL
    drift from two displaced grand mobilities (stepper.f90:650-688), and the
    Brownian stresslet SijN_B = RPQ^T UB in yeta_mu(1) and (4), with the z the
    reference's own generator drew (exported by the replay)."""
    from rsd_oracle import StepConfig
    L, dt, T, kbt = G[tag + "L_dt_T0_kbt"]
    conf, radii = G[tag + "conf0"].copy(), G[tag + "radii"]
    cfg = StepConfig(LB=np.full(3, L), kbt=kbt, is_period=is_period)
    for k in range(len(G[tag + "ntdem"])):
        conf, upos, yeta, nt, ex = oracle.stepper_stokesian(cfg, conf, radii, dt, T, z=G[tag + "z"][k])
        T += dt
        assert nt == G[tag + "ntdem"][k]
        assert np.max(np.abs(conf - G[tag + "conf"][k])) < 1e-10 * L
        assert relerr(G[tag + "upos"][k], upos) < 1e-10
        if is_period:
            assert relerr(G[tag + "sb"][k], ex["S_B"]) < 1e-10
            assert relerr(G[tag + "shyd"][k], ex["S_hyd"]) < 1e-10
            assert relerr(G[tag + "yeta"][k][[0, 3]], yeta[[0, 3]]) < 1e-10
        assert abs(yeta[2] - G[tag + "yeta"][k][2]) <= 1e-10 * abs(yeta[2])


# ------------------------------------------------------ explicit image sets
def test_explicit_image_cube_matches_the_reference(oracle):
    """NR = NI = 4 set explicitly reproduces the reference at the EWS_ERR that
    gives 4 and 4 by its own rule (lambda does not depend on EWS_ERR)."""
    conf, radii = G["img4_conf"], G["img4_radii"]
    L, T = G["img4_LT"]
    try:
        oracle.set_images(4, 4, 0)
        APP, APQ, AQQ, _ = oracle.per_grpery_inv_fri(conf, radii, _lat(oracle, T, [L, L, L], 1.0), 1e-4)
    finally:
        oracle.set_images()
    assert relerr(G["img4_APP"], APP) < TOL and relerr(G["img4_APQ"], APQ) < TOL and relerr(G["img4_AQQ"], AQQ) < TOL


def test_spherical_cutoff_properties(oracle):
    """A cutoff that holds the whole cube (kmax2 = 3 NI^2) is the cube; the
    configurations' set (nr 4, |n| <= 6) is converged: within 1e-8 of a much
    larger set (with lambda = sqrt(pi)/L the weights past |n| = 4 fall below
    exp(-16 pi)), while a cutoff at |n| <= 2 is visibly short."""
    conf, radii = G["img4_conf"], G["img4_radii"]
    L, T = G["img4_LT"]

    def grand(nr, ni, k2):
        oracle.set_images(nr, ni, k2)
        try:
            return oracle.per_grpery_inv_fri(conf, radii, _lat(oracle, T, [L, L, L], 1.0), 1e-4)[:3]
        finally:
            oracle.set_images()
    cube = grand(3, 3, 0)
    full = grand(3, 3, 27)
    for a, b in zip(cube, full):
        assert np.array_equal(a, b)
    conv = grand(5, 8, 0)
    cfg = grand(4, 6, 36)
    small = grand(4, 6, 4)
    for a, b, c in zip(conv, cfg, small):
        assert relerr(a, b) < 1e-8
        assert relerr(a, c) > 1e-7
