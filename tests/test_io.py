"""On-disk formats and units (CPU): the library's readers and writers against
the reference's own (initial_conf.f90, mymodules.f90:808-881), whose outputs
golden.npz holds (tests/golden/make_golden.py, io_* cases), and the drop-in
driver's refusal of settings outside the periodic FTS path."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

G = np.load(os.path.join(GOLDEN, "golden.npz"))
INT_FIELDS = ["correction_method", "fts_method", "filament_solve_implicit", "uselub", "use_dem", "usecollision",
              "use_dem_stepper", "use_gmres", "xshear", "is_period", "simple_period", "oscillation_shear", "fix",
              "is_case_lattice", "beginstep", "interval", "num_threads", "timestep", "collision_condition",
              "mu_calculate", "k_write", "boundary", "wall_method", "use_dlvo"]
REAL_FIELDS = (["radii_test", "radii_true", "g_gravity"] + [("u_bg", k) for k in range(3)] +
               [("omega_bg", k) for k in range(3)] + [("ei_bg", k) for k in range(6)] +
               ["mu_f", "pho_f", "re_number", "gamma_alpha", "youngs_module", "poisson_ratio", "ksp", "erest",
                "pho_par", "linerr", "t0", "end_time", "base_dt", "lj_eps", "lj_sigma", "lj_cut", "debye_length", "z0",
                "lambda_l", "permittivity", "potential", "ah", "surfaceenergy", "kbt"])
UNIT_FIELDS = ["lambda_time", "mu_f", "gravity", "end_time", "base_dt", "t0", "dt_dem0", "kbt", "frequency", "gamma"]


def _text(key):
    return bytes(G[key]).decode()


def _write(tmp_path, name, key):
    p = tmp_path / name
    p.write_text(_text(key))
    return p


@pytest.fixture(scope="module")
def lib():
    import rsd_amd
    return rsd_amd


@pytest.mark.parametrize("tag", ["fts", "brown"])
def test_read_control_matches_the_reference(lib, tmp_path, tag):
    c = lib.read_control(_write(tmp_path, "control_file.dat", f"io_control_{tag}_text"))
    ints, reals = G[f"io_control_{tag}_ints"], G[f"io_control_{tag}_reals"]
    for k, name in enumerate(INT_FIELDS):
        assert getattr(c, name) == ints[k], name
    for k, f in enumerate(REAL_FIELDS):
        v = getattr(c, f[0])[f[1]] if isinstance(f, tuple) else getattr(c, f)
        assert v == reals[k], f
    assert c.brown == int(reals[39])


@pytest.mark.parametrize("tag", ["fts", "brown"])
def test_units_match_the_reference(lib, tmp_path, tag):
    c = lib.read_control(_write(tmp_path, "control_file.dat", f"io_control_{tag}_text"))
    u = lib.nondimensionalise(c)
    want = G[f"io_units_{tag}"]
    for k, name in enumerate(UNIT_FIELDS):
        assert getattr(u, name) == pytest.approx(want[k], rel=1e-15, abs=0.0), name
    if tag == "fts":
        assert u.mu_f == pytest.approx(1.0 / np.pi, rel=1e-15)  # mymodules.f90:816,846


@pytest.mark.parametrize("tag", ["plain", "hot"])
def test_read_initial_config_matches_the_reference(lib, tmp_path, tag):
    conf, radii, lb, T, up, hot = lib.read_initial_config(_write(tmp_path, "initial_config.dat",
                                                                 f"io_config_{tag}_text"))
    want = G[f"io_config_{tag}_out"]
    n = len(radii)
    assert n == 6 and hot == (tag == "hot")
    got = np.concatenate([conf.ravel(), radii, lb, [T], up[:3 * n]])
    if tag == "plain":
        got[-3 * n:] = 0.0
    assert np.array_equal(got, want)


def test_writers_match_the_reference_text(lib, tmp_path):
    vtf, yeta = tmp_path / "output.VTF", tmp_path / "yeta_mu.txt"
    conf, radii, up = G["io_write_conf"], G["io_write_radii"], G["io_write_upos"]
    T0, T1 = G["io_write_T"]
    lib.write_vtf_frame(vtf, conf, T0, radii, up, append=False)
    lib.write_vtf_frame(vtf, conf * -0.5, T1, radii, up * 3, append=True)
    for k, row in enumerate(G["io_write_yrows"]):
        lib.write_yeta_mu(yeta, row[0], row[1:6], row[6], row[7], row[8], row[9], append=k > 0)
    assert vtf.read_text() == _text("io_vtf_text")
    assert yeta.read_text() == _text("io_yeta_text")


def test_hot_start_reads_what_vmd_write_writes(lib, tmp_path):
    """A VTF frame pasted under an initial_config.dat header with hot_start true
    gives back the frame's positions, radii and velocities to 15 digits."""
    from rsd_amd.configs import random_suspension
    conf, radii, L = random_suspension(7, 0.25, 9)
    up = np.concatenate([np.random.default_rng(1).normal(size=21), np.zeros(21)])
    vtf = tmp_path / "output.VTF"
    lib.write_vtf_frame(vtf, conf, 0.125, radii, up, append=False)
    hdr = f"0 0 0\n7\n0 0\n\n0 0\n\n0 0 0 0 0 0\n{L} {L} {L}\n.true.\n"
    (tmp_path / "initial_config.dat").write_text(hdr + vtf.read_text())
    c, r, lb, T, u, hot = lib.read_initial_config(tmp_path / "initial_config.dat")
    assert hot and T == 0.125 and np.allclose(lb, L, rtol=1e-15)
    assert np.allclose(c, conf, rtol=1e-14, atol=0) and np.array_equal(r, radii)
    assert np.allclose(u[:21], up[:21], rtol=1e-14, atol=0)


def test_driver_refuses_settings_outside_the_path(lib, tmp_path):
    text = _text("io_control_fts_text").replace(".false.                  !useGMRESmethod",
                                                ".true.                   !useGMRESmethod")
    (tmp_path / "control_file.dat").write_text(text.replace("2                        !correction_method",
                                                            "1                        !correction_method"))
    p = subprocess.run([lib.RUN_PATH], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "rsd_run:" in p.stderr


def test_live_reference_writer_text(lib, reference, tmp_path):
    """Fresh random rows through the reference's VMD_WRITE: the library writes the same text."""
    rng = np.random.default_rng(11)
    N = 5
    reference.setup(np.zeros((N, 3)), np.ones(N), np.full(3, 10.0))
    conf = rng.normal(size=(N, 3)) * 10.0 ** rng.integers(-20, 20, (N, 3))
    radii = rng.uniform(0.5, 2.0, N)
    up = np.concatenate([rng.normal(size=3 * N) * 10.0 ** rng.integers(-9, 9, 3 * N), np.zeros(3 * N)])
    y = rng.normal(size=5) * 10.0 ** rng.integers(-30, 30, 5)
    d = tmp_path / "ref"
    d.mkdir()
    vtf_ref, yeta_ref = reference.write_outputs(str(d), [(conf, 0.3, radii, up)], [(0.3, y, 12.5, 1.0, 2.0, 0.1)])
    lib.write_vtf_frame(tmp_path / "o.VTF", conf, 0.3, radii, up, append=False)
    lib.write_yeta_mu(tmp_path / "y.txt", 0.3, y, 12.5, 1.0, 2.0, 0.1, append=False)
    assert (tmp_path / "o.VTF").read_text() == vtf_ref
    assert (tmp_path / "y.txt").read_text() == yeta_ref
