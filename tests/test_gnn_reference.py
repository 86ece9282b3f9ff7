"""The float64 PyTorch statement of the learned preconditioner (the GPU
gradient check's reference) pinned on the CPU by closed forms: with the pair
weights driven to zero it is block-Jacobi; for two spheres with the weights
driven to one it is the exact inverse of K (loss zero); and its autograd
gradient agrees with a central difference."""
import numpy as np
import pytest

import gnn_reference as ref

N_PARAMS = 50629


def _system(oracle, conf, radii, L):
    lat = oracle.lattice(0.0, np.full(3, L), 1.0)
    APP, APQ, AQQ, _ = oracle.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    rfu, rfe, rse, _ = oracle.lubmxcalc(conf, radii, lat)
    M = np.block([[APP, APQ], [APQ.T, AQQ]])
    R = np.block([[rfu, rfe], [rfe.T, rse]])
    N = len(radii)
    d = conf[:, None] - conf[None]
    d -= L * np.round(d / L)
    r = np.linalg.norm(d, axis=2) + 1e9 * np.eye(N)
    cap = N
    nnb = np.zeros(N, np.int32)
    col = np.zeros((N, cap), np.int32)
    dist = np.zeros((N, cap))
    for g in range(N):
        nb = np.where(r[g] <= 4.0)[0]
        nnb[g] = len(nb)
        col[g, :len(nb)] = nb
        dist[g, :len(nb)] = r[g, nb]
    return M, R, nnb, col, dist


def test_two_spheres_exact_inverse(oracle):
    conf = np.array([[3.0, 3.0, 3.0], [5.1, 3.2, 2.9]])
    radii = np.ones(2)
    M, R, nnb, col, dist = _system(oracle, conf, radii, 12.0)
    Z = np.random.default_rng(0).standard_normal((2, 22))
    p = np.zeros(N_PARAMS)
    p[-1] = -40.0
    l_bj, _ = ref.loss(p, M, R, nnb, col, dist, radii, Z)
    p[-1] = 40.0
    l_ex, _ = ref.loss(p, M, R, nnb, col, dist, radii, Z)
    assert l_ex < 1e-24 and l_bj > 1e-6


def test_zero_model_is_block_jacobi_and_gradient_matches_differences(oracle):
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(10, 0.3, 4)
    M, R, nnb, col, dist = _system(oracle, conf, radii, L)
    n6 = 6 * len(radii)
    K = R @ M
    K[:n6, :n6] += np.eye(n6)
    K[n6:] = M[n6:]
    Z = np.random.default_rng(1).standard_normal((2, K.shape[0]))
    g, q, d = ref.edges(nnb, col)
    D, _, _ = ref.pair_pieces(K, len(radii), g, d)
    P = np.zeros_like(K)
    for i in range(len(radii)):
        P[np.ix_(ref.dofs(len(radii), i), ref.dofs(len(radii), i))] = D[i]
    r = K @ P @ Z.T - Z.T
    p0 = np.zeros(N_PARAMS)
    p0[-1] = -40.0
    l0, _ = ref.loss(p0, M, R, nnb, col, dist, radii, Z)
    assert l0 == pytest.approx(np.sum(r ** 2) / np.sum(Z ** 2), rel=1e-12)
    rng = np.random.default_rng(2)
    p = 0.1 * rng.standard_normal(N_PARAMS)
    l, grad = ref.loss(p, M, R, nnb, col, dist, radii, Z)
    for k in (5, 200, 192 + 16768 + 100, N_PARAMS - 1, N_PARAMS - 100):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (ref.loss(pp, M, R, nnb, col, dist, radii, Z)[0] - ref.loss(pm, M, R, nnb, col, dist, radii, Z)[0]) / (2 * h)
        assert fd == pytest.approx(grad[k], rel=1e-5, abs=1e-9 * abs(l))
