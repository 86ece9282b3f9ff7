"""The Fortran drop-in (regularized-stokesian-dynamics_amd/fortran/).

``oracle/_ref/libref_gpu.so`` is the reference's own Fortran (stepper.f90,
matrices.f90, ... as compiled by ``make -C oracle ref``) linked with the
drop-in in place of per_GRPY_YZD.f90 and hydro_lub.f90 and ahead of
matrices.f90's MATREV / MobToResist_FTS / arrangement, so the reference's
STEPPER_Stokesian runs its hot path on the MI355X library through
iso_c_binding.  Tolerances: 1e-12 relative for single evaluations, 1e-10 for
full steps (the hot path's FP64 parity bar).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import DATA, LUBDAT, ORACLE, relerr

LIB = os.path.join(ORACLE, "_ref", "libref_gpu.so")
needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="drop-in build absent (make -C oracle ref-gpu)")


@needs_lib
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is present")
def test_dropin_routes_matrev_to_the_library():
    """Without a device the drop-in's MATREV stops in rsd_device_init: the
    reference's call went to the library, and the failure is loud."""
    code = ("import sys, numpy as np; sys.path.insert(0, %r); from rsd_oracle import Reference; "
            "r = Reference(path=%r, lubdat_parent=%r); r.lib.ref_matrev(2, np.eye(2).ctypes.data_as("
            "__import__('ctypes').POINTER(__import__('ctypes').c_double)))" % (ORACLE, LIB, DATA))
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "rsd_device_init failed" in p.stdout + p.stderr


@pytest.fixture(scope="module")
def dropin():
    if not os.path.exists(LIB):
        pytest.skip("drop-in build absent (make -C oracle ref-gpu)")
    from rsd_oracle import Reference
    return Reference(path=LIB, lubdat_parent=DATA)


@pytest.mark.gpu
def test_dropin_grand_mobility_and_lubrication(dropin, oracle):
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(16, 0.35, seed=8)
    dropin.setup(conf, radii, np.full(3, L), gamma=1.0, linerr=1e-4)
    dropin.set_lattice_skew(0.3)
    APP, APQ, AQQ = dropin.per_grpery_inv_fri(conf, radii, 1e-4)
    lat = oracle.lattice(0.3, np.full(3, L), 1.0)
    oAPP, oAPQ, oAQQ, _ = oracle.per_grpery_inv_fri(conf, radii, lat, 1e-4)
    assert relerr(oAPP, APP) < 1e-12 and relerr(oAPQ, APQ) < 1e-12 and relerr(oAQQ, AQQ) < 1e-12
    g = dropin.get_lattice()  # CALC_LATTICE_INVERSE outputs written back into LATTICE_BASE
    assert (g["NR"], g["NI"]) == (lat.NR, lat.NI) and abs(g["lam"] - lat.lam) < 1e-15 and abs(g["phi"] - lat.phi) < 1e-15
    rfu, rfe, rse = dropin.lubmxcalc(conf)
    orfu, orfe, orse, _ = oracle.lubmxcalc(conf, radii, lat)
    assert relerr(orfu, rfu) < 1e-12 and relerr(orfe, rfe) < 1e-12 and relerr(orse, rse) < 1e-12


@pytest.mark.gpu
def test_dropin_reference_stepper(dropin, oracle):
    """Two steps of the reference's own STEPPER_Stokesian on the GPU path."""
    from rsd_oracle import StepConfig
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(16, 0.3, seed=21)
    dropin.setup(conf, radii, np.full(3, L), gamma=1.0, linerr=1e-4)
    cfg = StepConfig(LB=np.full(3, L))
    c_g, c_o, u_g, T, dt = conf.copy(), conf.copy(), None, 0.0, 1e-3
    for _ in range(2):
        c_g, u_g, y_g, nt_g = dropin.stepper_stokesian(c_g, radii, dt, T, u_g)
        c_o, u_o, y_o, nt_o, _ = oracle.stepper_stokesian(cfg, c_o, radii, dt, T)
        T += dt
        assert nt_g == nt_o
        assert np.max(np.abs(c_g - c_o)) < 1e-12 * L
        assert relerr(u_o, u_g) < 1e-10
        assert relerr(y_o[:3], y_g[:3]) < 1e-10
