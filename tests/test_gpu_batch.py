"""Batches of independent suspensions (rsd_batch): uneven sizes and volume
fractions, bidisperse radii, stepped together on concurrent host threads and
streams.  A batch must give, bit for bit, what each suspension gives stepped
alone; the width autotuner must leave every state as it found it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [(60, 0.15, 31), (150, 0.3, 32), (40, 0.45, 33), (220, 0.25, 34), (90, 0.4, 35), (120, 0.1, 36)]


def _configs():
    from rsd_amd.configs import bidisperse_radii, random_suspension
    return [random_suspension(n, phi, seed, radii=bidisperse_radii(n, seed)) for n, phi, seed in SPECS]


def _run_alone(rsd, cfgs, dt, steps):
    out = []
    for conf, radii, L in cfgs:
        sus = rsd.Suspension(conf, radii, rsd.default_config(L))
        T, ys = 0.0, []
        for _ in range(steps):
            y, info = sus.step(dt, T)
            assert info == 0
            ys.append(y)
            T += dt
        out.append((sus.state(), sus.last(), np.array(ys)))
        sus.close()
    return out


def test_batch_matches_each_suspension_alone(rsd):
    cfgs, dt, steps = _configs(), 5e-4, 3
    alone = _run_alone(rsd, cfgs, dt, steps)
    sus = [rsd.Suspension(conf, radii, rsd.default_config(L)) for conf, radii, L in cfgs]
    batch = rsd.Batch(sus, width=3)
    ys = []
    for k in range(steps):
        y, info = batch.step(dt, k * dt)
        assert np.all(info == 0)
        ys.append(y)
    ys = np.array(ys)
    for i, s in enumerate(sus):
        (c_a, u_a), (p_a, s_a, n_a), y_a = alone[i]
        c_b, u_b = s.state()
        p_b, s_b, n_b = s.last()
        assert np.array_equal(c_a, c_b) and np.array_equal(u_a, u_b), i
        assert np.array_equal(p_a, p_b) and np.array_equal(s_a, s_b) and n_a == n_b, i
        assert np.array_equal(y_a, ys[:, i, :]), i
    print(f"batch of {len(sus)} (N {[n for n, _, _ in SPECS]}): identical to each suspension alone over {steps} steps",
          flush=True)
    batch.close()
    for s in sus:
        s.close()


def test_batch_autotune_restores_the_states(rsd):
    cfgs = _configs()
    sus = [rsd.Suspension(conf, radii, rsd.default_config(L)) for conf, radii, L in cfgs]
    batch = rsd.Batch(sus)
    before = [s.state() for s in sus]
    best, ms = batch.autotune(5e-4, 0.0, widths=(1, 2, 6))
    print(f"autotuned width {best}; batch step ms by width {ms}", flush=True)
    assert best in (1, 2, 6) and batch.width() == best and all(v > 0 for v in ms.values())
    for s, (c0, u0) in zip(sus, before):
        c1, u1 = s.state()
        assert np.array_equal(c0, c1) and np.array_equal(u0, u1)
    batch.close()
    for s in sus:
        s.close()
