"""Seeded synthetic suspensions for the benchmark configurations.

Random sequential addition (RSA) with a minimum surface gap in a cubic
periodic box; above the RSA jamming density (about 0.38 for equal spheres)
the packing is grown at lower density and compressed in small steps, each
followed by deterministic pairwise overlap relaxation, until the target
volume fraction is reached with every gap >= ``gap``.

The reference reads its particles from ``initial_config.dat``
(initial_conf.f90:442-476); this module only produces those inputs.
"""
from __future__ import annotations

import numpy as np


def box_edge(radii, phi):
    """Cubic box edge L with sum(4/3 pi a^3) / L^3 = phi."""
    vol = np.sum(4.0 / 3.0 * np.pi * np.asarray(radii, dtype=np.float64) ** 3)
    return float((vol / phi) ** (1.0 / 3.0))


def _min_image(d, L):
    return d - L * np.round(d / L)


def _rsa(radii, L, gap, rng, max_tries=2_000_000):
    N = len(radii)
    pos = np.empty((N, 3))
    placed = 0
    tries = 0
    order = np.argsort(-radii, kind="stable")  # big spheres first
    out_r = radii[order]
    while placed < N:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("RSA did not converge; lower the starting volume fraction")
        p = rng.uniform(0.0, L, 3)
        if placed:
            d = _min_image(pos[:placed] - p, L)
            r = np.sqrt(np.einsum("ij,ij->i", d, d))
            if np.any(r < out_r[:placed] + out_r[placed] + gap):
                continue
        pos[placed] = p
        placed += 1
    inv = np.empty(N, dtype=int)
    inv[order] = np.arange(N)
    return pos[inv]


def _relax(pos, radii, L, gap, max_iter=2000):
    """Push overlapping pairs (surface gap < gap) apart along their line of centres."""
    contact = radii[:, None] + radii[None, :] + gap
    target = contact + 0.05 * gap  # push slightly past the limit so the test below is robust to rounding
    for _ in range(max_iter):
        d = _min_image(pos[:, None, :] - pos[None, :, :], L)
        r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
        np.fill_diagonal(r, np.inf)
        if not np.any(r < contact):
            return pos
        over = np.clip(target - r, 0.0, None)
        np.fill_diagonal(over, 0.0)
        w = (0.5 * over / r)[:, :, None]
        pos = pos + np.sum(w * d, axis=1)
        pos = np.mod(pos, L)
    raise RuntimeError("overlap relaxation did not converge")


def _mc_sweeps(pos, radii, L, gap, rng, sweeps, step=0.1):
    """Hard-sphere Monte Carlo at fixed density: random single-particle moves
    accepted when every surface gap stays >= gap.  Spreads out the pairs that
    compression leaves sitting exactly at the minimum gap."""
    N = len(radii)
    for _ in range(sweeps):
        for i in rng.permutation(N):
            trial = np.mod(pos[i] + rng.uniform(-step, step, 3), L)
            d = _min_image(pos - trial, L)
            r = np.sqrt(np.einsum("ij,ij->i", d, d))
            r[i] = np.inf
            if np.all(r >= radii + radii[i] + gap):
                pos[i] = trial
    return pos


def random_suspension(n, phi, seed, radii=1.0, gap=1e-3, rsa_phi=0.3, mc_sweeps=20):
    """Return (conf (n,3), radii (n,), L) for a seeded periodic suspension.

    ``radii`` is a scalar (monodisperse) or an array of length n.  Above
    ``rsa_phi`` the RSA packing is compressed to ``phi`` and then equilibrated
    with ``mc_sweeps`` hard-sphere Monte Carlo sweeps.
    """
    rng = np.random.default_rng(seed)
    radii = np.full(n, float(radii)) if np.isscalar(radii) else np.asarray(radii, dtype=np.float64).copy()
    L = box_edge(radii, phi)
    if phi <= rsa_phi:
        return _rsa(radii, L, gap, rng), radii, L
    L0 = box_edge(radii, rsa_phi)
    pos = _rsa(radii, L0, gap, rng)
    Lc = L0
    while Lc > L:
        Ln = max(L, Lc * 0.98)
        pos = pos * (Ln / Lc)
        Lc = Ln
        pos = _relax(pos, radii, Lc, gap)
    pos = _mc_sweeps(np.mod(pos, L), radii, L, gap, rng, mc_sweeps)
    return np.mod(pos, L), radii, L


def bidisperse_radii(n, seed, small=1.0, large=1.4):
    """50/50 bidisperse radii in a seeded random order."""
    rng = np.random.default_rng(seed)
    r = np.where(np.arange(n) < n // 2, small, large).astype(np.float64)
    rng.shuffle(r)
    return r
