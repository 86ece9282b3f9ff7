"""The learned preconditioner of the iterative route on the GPU, through the C
ABI: its inputs (the block-sparse lubrication against the dense lubmxcalc,
the grand mobility), its loss and hand-written backward pass against a float64
PyTorch statement of the same network (tests/gnn_reference.py), a closed form
(two spheres: the pair piece is the exact inverse), training, and its gain
over block-Jacobi on held-out suspensions with the committed trained model.
"""
import os

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _suspension(rsd, n, phi, seed):
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(n, phi, seed)
    sus = rsd.Suspension(conf, radii, rsd.default_config(L))
    return sus, conf, radii, L


def test_block_sparse_lubrication_matches_lubmxcalc(rsd):
    """The 11x11 neighbour blocks the iterative route multiplies by are
    hydro_lub.f90's lubmxcalc (rsd_lubmxcalc, pinned against the reference)
    and the system's M is PER_GRPERY_INV_FRI's."""
    sus, conf, radii, L = _suspension(rsd, 40, 0.4, 3)
    T = 0.2
    sus.prepare_iterative(T)
    M, R, nnb, col, dist, cap = sus.iterative_system()
    lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
    rfu, rfe, rse, _ = rsd.lubmxcalc(conf, radii, lat)
    Rd = np.block([[rfu, rfe], [rfe.T, rse]])
    APP, APQ, AQQ, _ = rsd.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    Md = np.block([[APP, APQ], [APQ.T, AQQ]])
    print(f"N=40: neighbours per particle {nnb.mean():.1f} (max {nnb.max()}, cap {cap})", flush=True)
    assert relerr(Rd, R) < 1e-12
    assert relerr(Md, M) < 1e-12
    for g in range(40):                       # symmetric neighbour lists, distances of the pairs
        for q in range(nnb[g]):
            d = col[g, q]
            assert g in col[d, :nnb[d]]
            assert dist[g, q] <= 4.0 + 1e-12
    sus.close()


def test_two_spheres_pair_piece_is_the_exact_inverse(rsd):
    """N = 2 with one pair: the pair weight driven to one (only the output
    bias set, to +40) makes P the inverse of K's whole 22x22 block, so K P = I
    and the loss vanishes; driven to zero (-40) P is block-Jacobi, whose loss
    does not."""
    conf = np.array([[3.0, 3.0, 3.0], [5.1, 3.2, 2.9]])
    radii = np.ones(2)
    sus = rsd.Suspension(conf, radii, rsd.default_config(12.0))
    sus.prepare_iterative(0.0)
    gnn = rsd.Gnn(seed=1)
    p = np.zeros(rsd.GNN_PARAMS, np.float32)
    p[-1] = -40.0
    gnn.set_params(p)
    loss_bj, _, _ = sus.gnn_loss(gnn, nz=2, grad=False)
    p[-1] = 40.0
    gnn.set_params(p)
    loss_ex, _, _ = sus.gnn_loss(gnn, nz=2, grad=False)
    print(f"two spheres: block-Jacobi loss {loss_bj:.3e}, exact pair inverse {loss_ex:.3e}", flush=True)
    assert loss_ex < 1e-20 and loss_bj > 1e-6
    gnn.close()
    sus.close()


def test_gnn_loss_and_gradient_match_torch(rsd):
    """Loss and all 50629 parameter gradients of the hand-written kernels
    against float64 autograd, with random weights everywhere (output layer
    included, so that every layer's gradient is nonzero)."""
    import gnn_reference as ref
    sus, conf, radii, L = _suspension(rsd, 24, 0.4, 5)
    sus.prepare_iterative(0.1)
    M, R, nnb, col, dist, cap = sus.iterative_system()
    gnn = rsd.Gnn(seed=3)
    p = gnn.params()
    rng = np.random.default_rng(4)
    p[-133:] = (0.05 * rng.standard_normal(133)).astype(np.float32)
    gnn.set_params(p)
    loss_g, grad_g, Z = sus.gnn_loss(gnn, nz=3, seed=11)
    loss_t, grad_t = ref.loss(p.astype(np.float64), M, R, nnb, col, dist, radii, Z)
    groups = {"W0,b0": slice(0, 192), "layer 0": slice(192, 192 + 16768), "layer 1": slice(192 + 16768, 192 + 2 * 16768),
              "layer 2": slice(192 + 2 * 16768, 192 + 3 * 16768), "Wo,bo": slice(192 + 3 * 16768, None)}
    errs = {k: relerr(grad_t[v], grad_g[v]) for k, v in groups.items()}
    print(f"GNN loss GPU {loss_g:.6e} torch {loss_t:.6e}; gradient relative errors {errs}", flush=True)
    assert abs(loss_g - loss_t) <= 1e-5 * loss_t
    for k, v in groups.items():
        assert np.linalg.norm(grad_t[v]) > 0.0, k
        assert errs[k] < 1e-3, k
    gnn.close()
    sus.close()


def test_training_lowers_the_loss(rsd):
    sus, _, _, _ = _suspension(rsd, 64, 0.45, 8)
    sus.prepare_iterative(0.0)
    gnn = rsd.Gnn(seed=2)
    l0, _, _ = sus.gnn_loss(gnn, nz=4, seed=99, grad=False)
    losses = sus.gnn_train(gnn, 150, nz=4, lr=2e-3, seed=5)
    l1, _, _ = sus.gnn_loss(gnn, nz=4, seed=99, grad=False)
    print(f"training on N=64: loss {l0:.3e} -> {l1:.3e} (first/last minibatch {losses[0]:.3e} {losses[-1]:.3e})",
          flush=True)
    assert np.all(np.isfinite(losses)) and l1 < 0.7 * l0
    gnn.close()
    sus.close()


def test_learned_preconditioner_beats_block_jacobi(rsd):
    """The committed model (data/gnn_params.npy, trained on the GPU by
    rsd_amd.gnn_train on suspensions of N = 256, phi in [0.35, 0.5]) on
    held-out suspensions (seeds outside the training set), GMRES iterations to
    a 1e-10 relative residual: alone, at most 0.3 of block-Jacobi's; as the
    mixed-precision inner preconditioner, fewer outer iterations than the
    mixed preconditioner with block-Jacobi inside, also at phi = 0.47 where
    block-Jacobi's restarted GMRES stalls."""
    path = rsd.GNN_PARAMS_PATH
    if not os.path.exists(path):
        pytest.skip("no trained model committed (python -m rsd_amd.gnn_train)")
    gnn = rsd.Gnn.load(path)
    for seed, phi in ((90001, 0.38), (90002, 0.42), (90003, 0.47)):
        sus, _, _, _ = _suspension(rsd, 256, phi, seed)
        sus.set_gnn(gnn)
        sus.prepare_iterative(0.0)
        it = {pc: sus.iterative_solve(pc, tol=1e-10, max_iter=2000)[0] for pc in (1, 3, 2, 4)}
        print(f"held-out N=256 phi={phi}: GMRES iterations block-Jacobi {it[1]}, learned {it[3]}; "
              f"mixed {it[2]}, mixed+learned {it[4]}", flush=True)
        if phi < 0.45:
            assert it[3] <= 0.3 * it[1]
        assert it[4] < it[2]
        sus.close()
    gnn.close()


def test_learned_preconditioner_gain_at_n1000(rsd):
    """Config 4's evaluation size: three held-out suspensions of N=1000 (phi
    0.38, 0.42, 0.46; tests/golden/config4_suspensions.npz).  To a 1e-10
    relative residual the learned preconditioner takes fewer GMRES iterations
    and less time than block-Jacobi (which does not converge in 2000 from phi
    0.42 on; neither does the learned one alone at 0.46), and inside the FP32
    inner solves fewer outer iterations and less time than block-Jacobi there,
    converging at every phi."""
    path = os.path.join(os.path.dirname(__file__), "golden", "config4_suspensions.npz")
    if not os.path.exists(path) or not os.path.exists(rsd.GNN_PARAMS_PATH):
        pytest.skip("config4_suspensions.npz or the trained model missing")
    F = np.load(path)
    gnn = rsd.Gnn.load(rsd.GNN_PARAMS_PATH)
    for k in range(3):
        L, phi, _ = F[f"L_phi_seed{k}"]
        sus = rsd.Suspension(F[f"conf{k}"], F[f"radii{k}"], rsd.default_config(float(L)))
        sus.set_gnn(gnn)
        sus.prepare_iterative(0.0)
        r = {}
        for pc in (1, 3, 2, 4):
            sus.iterative_solve(pc, tol=1e-10, max_iter=2000)
            r[pc] = sus.iterative_solve(pc, tol=1e-10, max_iter=2000)
        print(f"N=1000 phi={phi:.2f}: (iterations, relres, ms) block-Jacobi {r[1]}, learned {r[3]}; "
              f"mixed {r[2]}, mixed+learned {r[4]}", flush=True)
        assert r[4][1] <= 1e-10 and r[2][1] <= 1e-10
        assert r[4][0] < r[2][0] and r[4][2] < r[2][2]
        if phi < 0.45:
            assert r[3][1] <= 1e-10
            assert r[3][0] < r[1][0] and r[3][2] < r[1][2]
        sus.close()
    gnn.close()
