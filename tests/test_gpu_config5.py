"""Config 5: the full path at N=4096 (bidisperse radii 1 and 1.4 at phi=0.5,
Lees-Edwards shear), 11N = 45056: the FP64 grand mobility alone is 16 GB.
The reference's dense chain is out of the host's reach at this size.  The
configuration's path is the mixed-precision iterative route with the learned
preconditioner; it is held to the same route solved in FP64 throughout
(learned preconditioner, no FP32 inner solves), two independent solver
configurations, from the same suspension (tests/golden/config5_suspension.npz).
"""
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu


def test_config5_mixed_learned_route_against_fp64_direct(rsd):
    path = os.path.join(GOLDEN, "config5_suspension.npz")
    if not os.path.exists(path):
        pytest.skip("config5_suspension.npz not generated (python tests/golden/make_fixtures.py config5)")
    if not os.path.exists(rsd.GNN_PARAMS_PATH):
        pytest.skip("no trained model committed")
    F = np.load(path)
    L, dt = F["L_dt_phi_seed"][:2]
    gnn = rsd.Gnn.load(rsd.GNN_PARAMS_PATH)
    out = {}
    for route in ("direct", "mixed+learned"):
        sus = rsd.Suspension(F["conf0"], F["radii"], rsd.default_config(L))
        if route != "direct":
            sus.set_resistance_route(sus.RESIST_ITERATIVE)
            sus.set_iterative(precond=sus.PRECOND_MIXED_LEARNED, tol=1e-12, max_iter=2000)
            sus.set_gnn(gnn)
        rows, T = [], 0.0
        for k in range(2):
            t0 = time.perf_counter()
            y, info = sus.step(dt, T)
            ms = 1e3 * (time.perf_counter() - t0)
            up, s, nt = sus.last()
            st = sus.iterative_stats() if route != "direct" else sus.step_stats()
            print(f"config 5 (N=4096) {route} step {k}: {ms:.0f} ms, INFO={info}, NT_DEM={nt}, "
                  f"{'GMRES (iterations, relres, fallback)' if route != 'direct' else 'CG (iterations, fallbacks, redo)'}"
                  f" {st}", flush=True)
            assert info == 0
            if route != "direct":
                assert not st[2] and st[1] <= 1e-10
            rows.append((up.copy(), s.copy(), y.copy(), sus.state()[0].copy()))
            T += dt
        out[route] = rows
        sus.close()
    for k in range(2):
        (u1, s1, y1, c1), (u2, s2, y2, c2) = out["direct"][k], out["mixed+learned"][k]
        e = {"U_par": relerr(u1, u2), "SijN_hyd": relerr(s1, s2), "yeta": relerr(y1[:3], y2[:3]),
             "conf": float(np.max(np.abs(c1 - c2)) / L)}
        print(f"config 5 step {k}: mixed+learned vs FP64 direct, relative differences {e}", flush=True)
        assert max(e.values()) < 1e-10
    gnn.close()
