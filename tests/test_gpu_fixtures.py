"""The HIP path against the reference at the benchmark configurations' sizes
(tests/golden/make_fixtures.py: the flang-built reference's own outputs).

config 1: N=64, phi=0.2, 100 steps of dt=1e-3 under shear (collision
sub-steps from step 19 on): U_par, SijN_hyd and yeta_mu every step, positions
every 10 steps, all held to 1e-10 relative, the FP64 bar.  The reference's
own -O0/-O2 builds disagree by at most 7.5e-12 on the same steps (the spread
stored beside the outputs), so the bar is two orders above its floor.
config 2: N=512, phi=0.4 (the benchmark suspension), its first step, and its
200 steps until the overlapping spheres make the grand mobility indefinite,
where the step's INFO must be LAPACK's for that matrix.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu


def _load(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated (python tests/golden/make_fixtures.py)")
    return np.load(path)


@pytest.mark.parametrize("route,mode", [(1, 1), (0, 0)])
def test_config1_trajectory(rsd, route, mode):
    F = _load("config1_traj.npz")
    L, dt = F["L_dt_phi_seed"][:2]
    sus = rsd.Suspension(F["conf0"], F["radii"], rsd.default_config(L))
    sus.set_resistance_route(route)
    sus.set_substep_solver(mode)
    worst = {"upar": 0.0, "shyd": 0.0, "yeta": 0.0, "conf": 0.0}
    T = 0.0
    for k in range(len(F["ntdem"])):
        y, info = sus.step(dt, T)
        up, s, nt = sus.last()
        assert info == 0 and nt == F["ntdem"][k], k
        e = {"upar": relerr(F["upar"][k], up), "shyd": relerr(F["shyd"][k], s),
             "yeta": relerr(F["yeta"][k][:3], y[:3])}
        if k % 10 == 9:
            e["conf"] = float(np.max(np.abs(sus.state()[0] - F["conf_every10"][k // 10])) / L)
        for q, v in e.items():
            worst[q] = max(worst[q], v)
            assert v < 1e-10, (k, q, v)
        T += dt
    print(f"config 1, route {route} solver {mode}: worst relative errors over 100 steps {worst}; "
          f"reference -O0/-O2 spread: upar {F['spread_upar'].max():.1e} shyd {F['spread_shyd'].max():.1e}",
          flush=True)
    sus.close()


def test_config2_first_step(rsd):
    F = _load("config2_step.npz")
    L, dt = F["L_dt_phi_seed"][:2]
    sus = rsd.Suspension(F["conf0"], F["radii"], rsd.default_config(L))
    y, info = sus.step(dt, 0.0)
    c, u = sus.state()
    up, s, nt = sus.last()
    e = {"conf": relerr(F["conf"][0], c), "upos": relerr(F["upos"][0], u), "upar": relerr(F["upar"][0], up),
         "shyd": relerr(F["shyd"][0], s), "yeta": relerr(F["yeta"][0][:3], y[:3])}
    print(f"config 2 first step (N=512): relative errors {e}", flush=True)
    assert info == 0 and nt == F["ntdem"][0]
    assert max(e.values()) < 1e-10
    sus.close()


def test_config2_two_hundred_steps(rsd):
    """Config 2's 200 steps of dt=5e-4.  The reference has no contact force
    (source_inter_F is commented out, stepper.f90:96-97), so colliding spheres
    keep overlapping, and once the regularised grand mobility of the deepening
    overlaps is no longer positive definite, MATREV's DPOTRF fails (the
    reference prints its INFO and carries on with meaningless numbers).  The
    step must then report exactly the INFO LAPACK's DPOTRF('U') gives for the
    grand mobility of that configuration; every step before it factors, with
    finite, bounded viscosities."""
    from scipy.linalg import lapack
    F = _load("config2_step.npz")
    L, dt = F["L_dt_phi_seed"][:2]
    radii = F["radii"]
    sus = rsd.Suspension(F["conf0"], radii, rsd.default_config(L))
    T, etas, nts, failed = 0.0, [], [], None
    for k in range(200):
        conf = sus.state()[0]
        y, info = sus.step(dt, T)
        if info != 0:
            failed = (k, info, conf, T)
            break
        it, fb, redo = sus.step_stats()
        assert redo == 0 and fb == 0, (k, redo, fb)
        etas.append(y[0])
        nts.append(sus.last_nt())
        T += dt
    sus.close()
    print(f"config 2: {len(etas)} steps factored; yeta_mu(1) from {etas[0]:.4f} to {etas[-1]:.4f} "
          f"(min {min(etas):.4f}, max {max(etas):.4f}); steps with collision sub-steps {sum(n > 1 for n in nts)}",
          flush=True)
    assert np.all(np.isfinite(etas)) and 0.0 < min(etas) and max(etas) < 100.0
    if failed is not None:
        k, info, conf, T = failed
        lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
        APP, APQ, AQQ, nov = rsd.per_grpery_inv_fri(conf, radii, lat, 1e-4)
        M = np.block([[APP, APQ], [APQ.T, AQQ]])
        _, lapack_info = lapack.dpotrf(np.asfortranarray(M), lower=0, clean=0)
        d = conf[:, None, :] - conf[None, :, :]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=2) + 10.0 * np.eye(len(radii))
        print(f"config 2: step {k} (t={T:.4f}) reports INFO {info}; DPOTRF('U') of its grand mobility: "
              f"{lapack_info}; overlapping pairs {int(np.sum(r < 2.0)) // 2}, deepest gap {r.min() - 2.0:.3e}",
              flush=True)
        assert info == lapack_info > 0
