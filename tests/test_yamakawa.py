"""correction_method = 1: the Rotne-Prager-Yamakawa blocks of overlapping pairs
(YAMAKAWA_IJ, GRPY_YZD.f90:326-786) and the periodic sum's overlap correction
(overlap_Yamakawa_CORRECTION_IJ, GRPY_YZD.f90:789-811), on the CPU.

The oracle's literal restatement is pinned three ways: against the reference
itself (the flang build's YAMAKAWA_IJ and overlap_Yamakawa_CORRECTION_IJ, and
its PER_GRPERY_INV_FRI with correction_method = 1); against what the
mathematics fixes (the blocks are continuous where one sphere leaves the
other, r = |a_i - a_j|; inside, TT and RR are the larger sphere's own
mobilities 1/(6 a) and 1/(8 a^3); at contact, r = a_i + a_j, the translation
and rotation blocks meet the free RPY's); and against the pair symmetry
TT(R, a_i, a_j) = TT(-R, a_j, a_i).
"""
import numpy as np
import pytest

from conftest import relerr

NAMES = ("TT", "RR", "TR", "QQ", "GQ", "HQ")


def _pairs(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ai, aj = rng.uniform(0.6, 1.8, 2)
        v = rng.normal(size=3)
        out.append((rng.uniform(0.02, 0.999) * (ai + aj) * v / np.linalg.norm(v), ai, aj))
    out += [(np.array([0.2, -0.1, 0.15]), 1.0, 1.5), (np.array([0.2, -0.1, 0.15]), 1.5, 1.0),
            (np.array([1.2, 0.4, -0.3]), 1.0, 1.0)]
    return out


def test_yamakawa_blocks_match_the_reference(oracle, reference):
    for R, ai, aj in _pairs(7, 200):
        for mine, ref in ((oracle.rp_yamakawa, reference.rp_yamakawa),
                          (oracle.overlap_yamakawa_correction, reference.overlap_yamakawa_correction)):
            for name, a, b in zip(NAMES, mine(R, ai, aj), ref(R, ai, aj)):
                assert np.max(np.abs(a - b)) <= 1e-14 * max(1.0, np.max(np.abs(b))), (name, R, ai, aj)


def test_yamakawa_grand_mobility_matches_the_reference(oracle, reference):
    from rsd_amd.configs import bidisperse_radii, random_suspension
    conf, radii, L = random_suspension(10, 0.2, 5, radii=bidisperse_radii(10, 5))
    conf[1] = conf[0] + np.array([0.9, 0.3, -0.2])     # overlapping pairs
    conf[2] = conf[3] + np.array([-1.5, 0.1, 0.0])
    LB = np.full(3, L)
    reference.setup(conf, radii, LB, correction_method=1)
    reference.set_lattice_skew(0.0)
    reference.calc_lattice_inverse(1e-4)
    A = reference.per_grpery_inv_fri(conf, radii, 1e-4)
    B = oracle.per_grpery_inv_fri(conf, radii, oracle.lattice(0.0, LB, 1.0), 1e-4, correction_method=1)
    C2 = oracle.per_grpery_inv_fri(conf, radii, oracle.lattice(0.0, LB, 1.0), 1e-4, correction_method=2)
    assert B[3] >= 2
    for x, y, z in zip(A[:3], B[:3], C2[:3]):
        assert relerr(x, y) < 1e-13
        assert relerr(x, z) > 1e-6          # the two corrections differ: the test sees the method


def test_yamakawa_continuity_and_limits(oracle):
    t = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    for ai, aj in ((1.0, 1.4), (1.4, 1.0), (0.7, 1.6)):
        d0 = abs(ai - aj)
        lo, hi = oracle.rp_yamakawa(t * d0 * (1 - 1e-10), ai, aj), oracle.rp_yamakawa(t * d0 * (1 + 1e-10), ai, aj)
        for name, a, b in zip(NAMES, lo, hi):
            assert np.max(np.abs(a - b)) < 1e-9, (name, ai, aj)
        TT, RR = oracle.rp_yamakawa(t * 0.5 * d0, ai, aj)[:2]
        am = max(ai, aj)
        assert np.allclose(TT, np.eye(3) / (6 * am), rtol=0, atol=1e-15)
        assert np.allclose(RR, np.eye(3) / (8 * am ** 3), rtol=0, atol=1e-15)
    for ai, aj in ((1.0, 1.0), (1.0, 1.4), (1.4, 0.8)):
        c = oracle.overlap_yamakawa_correction(t * (ai + aj), ai, aj)
        for name, x in zip(NAMES[:3], c[:3]):
            assert np.max(np.abs(x)) < 1e-15, name


def test_yamakawa_pair_symmetry(oracle):
    for R, ai, aj in _pairs(8, 50):
        a = oracle.rp_yamakawa(R, ai, aj)
        b = oracle.rp_yamakawa(-R, aj, ai)
        assert np.allclose(a[0], b[0].T, rtol=0, atol=1e-14)   # TT_ij = TT_ji^T
        assert np.allclose(a[1], b[1].T, rtol=0, atol=1e-14)   # RR
