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


def test_config5_mixed_learned_route(rsd):
    path = os.path.join(GOLDEN, "config5_suspension.npz")
    if not os.path.exists(path):
        pytest.skip("config5_suspension.npz not generated (python tests/golden/make_fixtures.py config5)")
    if not os.path.exists(rsd.GNN_PARAMS_PATH):
        pytest.skip("no trained model committed")
    F = np.load(path)
    L, dt = F["L_dt_phi_seed"][:2]
    gnn = rsd.Gnn.load(rsd.GNN_PARAMS_PATH)
    out = {}
    for precond in (4, 2):
        sus = rsd.Suspension(F["conf0"], F["radii"], rsd.default_config(L))
        sus.set_resistance_route(sus.RESIST_ITERATIVE)
        sus.set_iterative(precond=precond, tol=1e-12, max_iter=2000)
        sus.set_gnn(gnn)
        rows, T = [], 0.0
        for k in range(2):
            t0 = time.perf_counter()
            y, info = sus.step(dt, T)
            ms = 1e3 * (time.perf_counter() - t0)
            up, s, nt = sus.last()
            its, rr, fb = sus.iterative_stats()
            print(f"config 5 (N=4096) precond {precond} step {k}: {ms:.0f} ms, INFO={info}, NT_DEM={nt}, "
                  f"GMRES iterations {its}, relres {rr:.1e}, fallback {fb}", flush=True)
            assert info == 0 and not fb and rr <= 1e-10
            rows.append((up.copy(), s.copy(), y.copy(), sus.state()[0].copy()))
            T += dt
        out[precond] = rows
        sus.close()
    for k in range(2):
        (u1, s1, y1, c1), (u2, s2, y2, c2) = out[4][k], out[2][k]
        e = {"U_par": relerr(u2, u1), "SijN_hyd": relerr(s2, s1), "yeta": relerr(y2[:3], y1[:3]),
             "conf": float(np.max(np.abs(c1 - c2)) / L)}
        print(f"config 5 step {k}: mixed+learned vs mixed (block-Jacobi inside) relative differences {e}",
              flush=True)
        assert max(e.values()) < 1e-10
    gnn.close()
