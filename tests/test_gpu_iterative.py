"""The iterative resistance route (RSD_RESIST_ITERATIVE) against the
reference's outputs, through the C ABI.

The route solves each sub-step's grand resistance problem in the mobility
form the reference's hydro_SD_FTS hands to GMRES (stepper.f90:476-546): one
system K x = [Fe; -EIN], K = [(I + R_lub M)_P ; M_Q], by flexible GMRES in
FP64 with the lubrication in 11x11 neighbour blocks, preconditioned by K's
diagonal blocks (block-Jacobi) or by FP32 inner GMRES iterations (mixed).  The
reference's own GMRES stops at ||residual|| < LINERR 1e-2; the route runs to a
1e-12 relative residual (or to FP64's floor within 1e-10, where dense
suspensions' lubrication puts it) so that its outputs meet the FP64 bar, 1e-10
relative against the reference's direct path (golden.npz, config1_traj.npz,
config2_step.npz).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(GOLDEN, "golden.npz"))


def _fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated (python tests/golden/make_fixtures.py)")
    return np.load(path)


@pytest.mark.parametrize("tag,precond", [("trajle_", 2), ("trajle_", 1), ("trajnb_", 2), ("traj_", 2),
                                         ("traj_", 0)])
def test_iterative_trajectory_golden(rsd, tag, precond):
    """Whole steps on the iterative route against the reference (the
    Lees-Edwards wraps, boundary particles and collision sub-steps of the
    direct route's golden trajectories)."""
    if tag == "traj_":
        L, dt = G["traj_L_dt"]
        T0, nb = 0.0, 0
    else:
        L, dt, T0, nb = G[tag + "L_dt_T0_nb"]
        nb = int(nb)
    sus = rsd.Suspension(G[tag + "conf0"], G[tag + "radii"], rsd.default_config(L, nb=nb))
    sus.set_resistance_route(sus.RESIST_ITERATIVE)
    sus.set_iterative(precond=precond, tol=1e-12, max_iter=4000)
    T = T0
    for k in range(len(G[tag + "ntdem"])):
        y, info = sus.step(dt, T)
        c, u = sus.state()
        up, s, nt = sus.last()
        its, rr, fb = sus.iterative_stats()
        print(f"iterative {tag} precond {precond} step {k}: NT_DEM={nt} GMRES iterations {its} "
              f"relres {rr:.1e} fallback {fb}", flush=True)
        assert info == 0 and nt == G[tag + "ntdem"][k]
        assert not fb and 0 < its and rr <= 1e-10
        assert np.max(np.abs(c - G[tag + "conf"][k])) < 1e-10 * L
        assert relerr(G[tag + "upos"][k], u) < 1e-10
        if tag != "traj_":
            assert relerr(G[tag + "upar"][k], up) < 1e-10
            assert relerr(G[tag + "shyd"][k], s) < 1e-10
        assert relerr(G[tag + "yeta"][k][:3], y[:3]) < 1e-10
        T += dt
    sus.close()


def test_iterative_config1(rsd):
    """N=64, phi=0.2: the first 30 steps (collision sub-steps included from
    step 19) on the mixed-precision iterative route, 1e-10 against the
    reference."""
    F = _fixture("config1_traj.npz")
    L, dt = F["L_dt_phi_seed"][:2]
    sus = rsd.Suspension(F["conf0"], F["radii"], rsd.default_config(L))
    sus.set_resistance_route(sus.RESIST_ITERATIVE)
    worst, T, total = {"upar": 0.0, "shyd": 0.0, "yeta": 0.0}, 0.0, 0
    for k in range(30):
        y, info = sus.step(dt, T)
        up, s, nt = sus.last()
        its, rr, fb = sus.iterative_stats()
        total += its
        assert info == 0 and nt == F["ntdem"][k] and not fb, k
        e = {"upar": relerr(F["upar"][k], up), "shyd": relerr(F["shyd"][k], s),
             "yeta": relerr(F["yeta"][k][:3], y[:3])}
        for q, v in e.items():
            worst[q] = max(worst[q], v)
            assert v < 1e-10, (k, q, v)
        T += dt
    print(f"config 1 iterative (mixed): worst relative errors over 30 steps {worst}; "
          f"{total / 30:.1f} outer GMRES iterations a step", flush=True)
    sus.close()


@pytest.mark.parametrize("precond", [2, 1])
def test_iterative_config2_first_step(rsd, precond):
    """N=512, phi=0.4 (the benchmark suspension): the first step on the
    iterative route, every output within 1e-10 of the reference."""
    F = _fixture("config2_step.npz")
    L, dt = F["L_dt_phi_seed"][:2]
    sus = rsd.Suspension(F["conf0"], F["radii"], rsd.default_config(L))
    sus.set_resistance_route(sus.RESIST_ITERATIVE)
    sus.set_iterative(precond=precond, tol=1e-12, max_iter=4000)
    y, info = sus.step(dt, 0.0)
    c, u = sus.state()
    up, s, nt = sus.last()
    its, rr, fb = sus.iterative_stats()
    e = {"conf": relerr(F["conf"][0], c), "upos": relerr(F["upos"][0], u), "upar": relerr(F["upar"][0], up),
         "shyd": relerr(F["shyd"][0], s), "yeta": relerr(F["yeta"][0][:3], y[:3])}
    print(f"config 2 first step, iterative precond {precond}: {its} GMRES iterations, relres {rr:.1e}, "
          f"relative errors {e}", flush=True)
    assert info == 0 and nt == F["ntdem"][0] and not fb
    assert max(e.values()) < 1e-10
    sus.close()


def test_iterative_gravity_matches_direct(rsd):
    """With gravity the route solves twice at the last sub-step (Fe for U,
    Fe = 0 for SijN_hyd, the stresslet the direct path's mu EIN gives)."""
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(40, 0.3, 7)
    out = []
    for route in (1, 2):
        sus = rsd.Suspension(conf, radii, rsd.default_config(L, gravity=0.7))
        sus.set_resistance_route(route)
        T = 0.0
        for k in range(3):
            y, info = sus.step(1e-3, T)
            assert info == 0
            T += 1e-3
        c, u = sus.state()
        up, s, nt = sus.last()
        if route == 2:
            assert not sus.iterative_stats()[2]
        out.append((c, u, up, s, y))
        sus.close()
    (c0, u0, p0, s0, y0), (c1, u1, p1, s1, y1) = out
    assert np.max(np.abs(c0 - c1)) < 1e-10 * L
    for a, b in ((u0, u1), (p0, p1), (s0, s1), (y0[:3], y1[:3])):
        assert relerr(a, b) < 1e-10


def test_iterative_fallback_redoes_the_step_directly(rsd):
    """A solve that does not converge (one GMRES iteration allowed) redoes the
    step on the Schur route from the step's start: the same outputs as a
    Schur step, and the stats say so."""
    L, dt = G["traj_L_dt"]
    conf, radii = G["traj_conf0"], G["traj_radii"]
    res = []
    for route in (2, 1):
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        sus.set_resistance_route(route)
        if route == 2:
            sus.set_iterative(precond=1, tol=1e-12, max_iter=1)
        y, info = sus.step(dt, 0.0)
        assert info == 0
        if route == 2:
            assert sus.iterative_stats()[2]
        res.append((sus.state(), sus.last(), y))
        sus.close()
    (sa, la, ya), (sb, lb, yb) = res
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])
    assert np.array_equal(la[0], lb[0]) and np.array_equal(la[1], lb[1])
    assert np.array_equal(ya, yb)


def test_inner_gmres_graph_replays_match_kernel_by_kernel(rsd):
    """The mixed preconditioners' inner FP32 GMRES captured as a HIP graph:
    replayed in three rounds of solves, with eager work between the rounds,
    every solve takes the same iterations to the same residual (bit for bit)
    as when its kernels are launched one by one."""
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(96, 0.4, 12)
    sus = rsd.Suspension(conf, radii, rsd.default_config(L))
    gnn = rsd.Gnn(seed=5)
    sus.set_gnn(gnn)
    sus.prepare_iterative(0.0)
    sus.set_iterative_graphs(False)
    eager = {pc: sus.iterative_solve(pc, tol=1e-10)[:2] for pc in (2, 4)}
    sus.set_iterative_graphs(True)
    for rnd in range(3):
        for pc in (2, 4):
            got = sus.iterative_solve(pc, tol=1e-10)[:2]
            print(f"round {rnd} precond {pc}: graph {got}, kernel by kernel {eager[pc]}", flush=True)
            assert got == eager[pc]
        sus.iterative_solve(1, tol=1e-8)            # eager work between the rounds
    sus.close()
    gnn.close()
