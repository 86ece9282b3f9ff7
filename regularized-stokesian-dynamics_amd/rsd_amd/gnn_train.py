"""Training and evaluation of the iterative route's learned preconditioner on
the GPU (the model and its loss: rsd_gnn.h).

    python -m rsd_amd.gnn_train --train 48 --n 256 --out data/gnn_params.npy

trains a fresh model on suspensions of n particles with volume fractions drawn
from [phi_lo, phi_hi] (seeded random sequential addition, as the benchmark
suspensions), a few Adam steps on each system in turn over several epochs,
and then reports GMRES iterations and solve times to a 1e-10 relative
residual on held-out suspensions (seeds from 90001 on) for block-Jacobi, the
learned preconditioner (trained, and untrained: every pair weight 1/2), and
both as the mixed-precision inner preconditioner.
Everything runs on the device; the host only generates the configurations.
"""
import argparse
import json
import time

import numpy as np

import rsd_amd
from rsd_amd.configs import random_suspension

PRECONDS = {"block_jacobi": 1, "learned": 3, "mixed": 2, "mixed_learned": 4}


def make_configs(count, n, phi_lo, phi_hi, seed0):
    rng = np.random.default_rng(seed0)
    out = []
    for k in range(count):
        phi = float(rng.uniform(phi_lo, phi_hi))
        conf, radii, L = random_suspension(n, phi, seed0 + k)
        out.append((conf, radii, L, phi))
    return out


def train(configs, epochs=3, iters=20, nz=4, lr=2e-3, seed=0, log=print):
    rsd_amd.init(0)
    gnn = rsd_amd.Gnn(seed=seed)
    sus = []
    for conf, radii, L, _ in configs:
        s = rsd_amd.Suspension(conf, radii, rsd_amd.default_config(L))
        s.prepare_iterative(0.0)
        sus.append(s)
    history = []
    for ep in range(epochs):
        t0 = time.time()
        first, last = [], []
        for k, s in enumerate(sus):
            losses = s.gnn_train(gnn, iters, nz=nz, lr=lr, seed=seed * 100003 + ep * 1009 + k)
            first.append(losses[0])
            last.append(losses[-1])
        history.append({"epoch": ep, "mean_first_loss": float(np.mean(first)), "mean_last_loss": float(np.mean(last)),
                        "seconds": time.time() - t0})
        log(f"epoch {ep}: mean loss {np.mean(first):.4e} -> {np.mean(last):.4e} ({time.time() - t0:.1f} s)")
    for s in sus:
        s.close()
    return gnn, history


def evaluate(gnn, configs, tol=1e-10, log=print):
    """GMRES iterations and solve times on each configuration, for each
    preconditioner, and for the untrained model (every pair weight 1/2)."""
    untrained = rsd_amd.Gnn(seed=0)
    rows = []
    for conf, radii, L, phi in configs:
        s = rsd_amd.Suspension(conf, radii, rsd_amd.default_config(L))
        s.prepare_iterative(0.0)
        row = {"phi": phi}
        for model, tag in ((gnn, ""), (untrained, "untrained_")):
            s.set_gnn(model)
            for name, pc in PRECONDS.items():
                if tag and pc not in (3, 4):
                    continue
                s.iterative_solve(pc, tol=tol)  # warm-up
                it, rr, ms = s.iterative_solve(pc, tol=tol)
                row[tag + name] = {"iterations": it, "relres": rr, "ms": ms}
        log(f"phi={phi:.3f}: " + ", ".join(f"{k} {v['iterations']} it {v['ms']:.1f} ms"
                                           for k, v in row.items() if k != "phi"))
        rows.append(row)
        s.close()
    untrained.close()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--train", type=int, default=48, help="training suspensions")
    ap.add_argument("--eval", type=int, default=6, help="held-out suspensions")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--phi-lo", type=float, default=0.35)
    ap.add_argument("--phi-hi", type=float, default=0.5)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--nz", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="where to save the parameters (.npy)")
    ap.add_argument("--report", default=None, help="where to write the training/evaluation record (.json)")
    a = ap.parse_args(argv)
    t0 = time.time()
    tr = make_configs(a.train, a.n, a.phi_lo, a.phi_hi, 1000)
    ev = make_configs(a.eval, a.n, a.phi_lo, a.phi_hi, 90001)
    print(f"{a.train} + {a.eval} configurations of N={a.n} in {time.time() - t0:.1f} s", flush=True)
    gnn, hist = train(tr, a.epochs, a.iters, a.nz, a.lr, a.seed, log=lambda m: print(m, flush=True))
    if a.out:
        gnn.save(a.out)
    rows = evaluate(gnn, ev, log=lambda m: print(m, flush=True))
    if a.report:
        with open(a.report, "w") as f:
            json.dump({"args": vars(a), "history": hist, "eval": rows}, f, indent=1)
    gnn.close()


if __name__ == "__main__":
    main()
