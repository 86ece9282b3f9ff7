#!/usr/bin/env python3
"""Benchmark: SD steps/s of the periodic FTS Stokesian-dynamics step on one MI355X.

Workload (the value): N=512 monodisperse spheres (a=1) at phi=0.4 in a cubic
box (L=17.502), simple shear gamma=1 with Lees-Edwards images, Ewald RPY grand
mobility with the reference's truncation rule (LINERR=1e-4), tabulated
lubrication, forward Euler dt=5e-4.  One step = STEPPER_Stokesian
(stepper.f90:33-156).  Inputs are resident on the device before the timed
region.  Two paths are timed on the same suspension:
  fp64_gpu       the direct FP64 route (Schur complement, CG sub-steps);
  mixed_learned  the iterative route: FGMRES in FP64 on the mobility form with
                 FP32 inner GMRES preconditioned by the learned pair correction.
``value`` is the faster one, the FP64 route (``--headline`` picks); the line
carries both and their ratio, the metric's comparison.

The first step of each path runs from the suspension of
tests/golden/config2_step.npz (the reference's own outputs for this step, from
the flang build of its sources) and is compared with it: ``parity`` holds the
relative errors of U_pos, U_par, SijN_hyd (the stresslets) and yeta_mu, beside
the reference's own -O0/-O2 spread.  Every step's INFO must be 0.

Beside it: config-mode Ewald (4 real images, |k| <= 6 2pi/L), the config-3
batch (64 suspensions, suspension-steps/s), the learned preconditioner's gain
at N=1000 (config 4) and two N=4096 steps (config 5) on the mixed+learned
route.  ``cpu_baseline`` times the reference Fortran (flang build in
oracle/_ref) on a bounded sample of the same step, and a minimal-algorithm CPU
variant beside it.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "regularized-stokesian-dynamics_amd"))

METRIC = "SD steps/sec at N=512 φ=0.4 (mixed+learned vs FP64 GPU); stresslet rel err"
PEAK_FP64 = 78.6e12   # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12     # B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--phi", type=float, default=0.4)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=2000)
    ap.add_argument("--cg-tol", type=float, default=None,
                    help="relative residual the CG sub-steps run to (default: the library's, 1e-14)")
    ap.add_argument("--substep-solver", choices=["direct", "pcg", "pcg-one-launch"], default="pcg",
                    help="collision sub-steps after the first: factor every sub-step (the reference's "
                         "arithmetic), CG with one launch per phase, or CG in one cooperative launch per solve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--config-mode-steps", type=int, default=5,
                    help="steps timed with the configurations' explicit Ewald image set (4 real images, |k| <= 6 2pi/L) "
                         "after the main run (0: skip)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event timing")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision learned iterative route")
    ap.add_argument("--gnn-params", default=os.path.join(ROOT, "regularized-stokesian-dynamics_amd", "data",
                                                         "gnn_params.npy"),
                    help="the learned preconditioner's parameters")
    ap.add_argument("--gmres-tol", type=float, default=1e-12)
    ap.add_argument("--inner-iters", type=int, default=20)
    ap.add_argument("--headline", choices=["mixed-learned", "fp64"], default="fp64",
                    help="which path's throughput is the line's value")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the mixed preconditioner's inner GMRES kernel by kernel (no HIP graph)")
    ap.add_argument("--no-batch", action="store_true", help="skip the config-3 batch line")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-config4", action="store_true", help="skip the learned preconditioner's N=1000 evaluation")
    ap.add_argument("--no-config5", action="store_true", help="skip the N=4096 line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the calibrated rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/ (None when absent)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.txt"))):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    return None


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(conf, radii, L, nt_dem):
    """Reference Fortran (oracle/_ref/libref.so, flang -O2 -fopenmp) timed on a
    bounded sample of the same step.  The reference's STEPPER_Stokesian runs
    the Ewald assembly once per step and lubmxcalc + STEPPER_NOGmres (whose
    MobToResist_FTS/MATREV chain is single-threaded vendored LAPACK) once per
    collision sub-step (stepper.f90:58-143), so the step estimate is
    assembly + NT_DEM x (lubrication + dense).  Sample: the full-N assembly
    and lubrication, and the dense chain at N/2 scaled by 8 (it is O(N^3)),
    each the median of three runs.

    ``minimal``: the same step with the least dense work the algorithm needs,
    on the host's multithreaded LAPACK (numpy/scipy's OpenBLAS): the
    reference's own assembly and lubrication, one Cholesky factorisation of
    the 11N grand mobility, W_PP from its Schur-complement factor, and per
    sub-step one Cholesky factorisation and solve of RPP (no explicit 11N
    inverse, no three 6N inverses)."""
    import numpy as np
    import scipy.linalg as sla
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        from rsd_oracle import Reference
        ref = Reference(lubdat_parent=os.path.join(ROOT, "regularized-stokesian-dynamics_amd", "data"))
    except Exception as e:  # reference not built here
        return {"value": None, "unit": "SD steps/s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {e}"}
    N = len(radii)
    LB = np.full(3, L)

    def median3(fn):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), r

    ref.setup(conf, radii, LB, gamma=1.0, linerr=1e-4)
    ref.set_lattice_skew(0.0)
    t_asm, (APP, APQ, AQQ) = median3(lambda: ref.per_grpery_inv_fri(conf, radii, 1e-4))
    t_lub, (rfu_f, rfe_f, rse_f) = median3(lambda: ref.lubmxcalc(conf))
    ns = N // 2
    sub, rsub = conf[:ns].copy(), radii[:ns].copy()
    ref.setup(sub, rsub, LB, gamma=1.0, linerr=1e-4)
    ref.set_lattice_skew(0.0)
    sAPP, sAPQ, sAQQ = ref.per_grpery_inv_fri(sub, rsub, 1e-4)
    rfu, rfe, rse = ref.lubmxcalc(sub)
    ein_s = ref.ein(ns)
    t_dense_s, _ = median3(lambda: ref.nogmres(sAPP, sAPQ, sAQQ, rfu, rfe, rse, np.zeros(6 * ns), ein_s))
    scale = (N / ns) ** 3
    t_dense = t_dense_s * scale
    total = t_asm + nt_dem * (t_lub + t_dense)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    # minimal-algorithm dense work at full N on the host's LAPACK
    n6, n5 = 6 * N, 5 * N
    ein = np.tile(ref.ein(1), N)
    t0 = time.perf_counter()
    Gq = np.block([[AQQ, APQ.T], [APQ, APP]])
    Lg = sla.cholesky(Gq, lower=True, check_finite=False)
    Ls = Lg[n5:, n5:]
    Lsi = sla.lapack.dtrtri(Ls, lower=1)[0]
    WPP = Lsi.T @ Lsi
    t_m_fact = time.perf_counter() - t0
    t0 = time.perf_counter()
    RPP = WPP + rfu_f
    cf = sla.cho_factor(RPP, lower=True, check_finite=False)
    sla.cho_solve(cf, np.ones(n6), check_finite=False)
    t_m_sub = time.perf_counter() - t0
    total_min = t_asm + t_m_fact + nt_dem * (t_lub + t_m_sub)
    blas_threads = os.environ.get("OPENBLAS_NUM_THREADS") or os.environ.get("OMP_NUM_THREADS") or str(os.cpu_count())
    return {"value": 1.0 / total, "unit": "SD steps/s", "cores": cores, "kind": "reference",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": (f"flang -O2 -fopenmp build of the reference, medians of 3 runs: PER_GRPERY_INV_FRI at N={N} "
                       f"({t_asm:.2f}s, {cores} OpenMP threads), lubmxcalc at N={N} ({t_lub:.2f}s), STEPPER_NOGmres "
                       f"dense chain (MobToResist_FTS + MATREV + arrangement, single-threaded vendored LAPACK as "
                       f"shipped) at N={ns} ({t_dense_s:.2f}s, x{scale:.0f} for N={N}); step = assembly + "
                       f"NT_DEM={nt_dem:g} x (lubrication + dense) = {total:.1f}s"),
            "stage_s": {"assembly": t_asm, "lubrication": t_lub, "dense": t_dense},
            "minimal": {"value": 1.0 / total_min, "unit": "SD steps/s", "kind": "port",
                        "blas_threads": blas_threads,
                        "sample": (f"the reference's assembly and lubrication as above, plus at N={N} on "
                                   f"numpy/scipy LAPACK: one Cholesky of the 11N grand mobility and W_PP from "
                                   f"its Schur factor ({t_m_fact:.2f}s), and per sub-step one Cholesky and "
                                   f"solve of RPP ({t_m_sub:.2f}s); step = {total_min:.2f}s")}}


def run_path(rsd_amd, np, sus, args, fix, label, dom_family=None):
    """One path of the step: its first step against the reference fixture,
    warmup, a per-family breakdown from untimed steps, then args.steps timed
    steps (only the dominant family's launches bracketed by events)."""
    T = 0.0
    bad_info, redos = [], 0
    it_stats = []

    def checked_step(T):
        y, info = sus.step(args.dt, T)   # returns after its stream has synchronised
        if info != 0:
            bad_info.append(info)
        return y, info

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))

    parity = None
    y1, _ = checked_step(T)
    T += args.dt
    if fix is not None:
        c1, u1 = sus.state()
        up1, s1, nt1 = sus.last()
        parity = {"reference": "tests/golden/config2_step.npz (flang -O2 build of the reference, first step)",
                  "nt_dem": [int(nt1), int(fix["ntdem"][0])],
                  "conf": rel(c1, fix["conf"][0]), "U_pos": rel(u1, fix["upos"][0]),
                  "U_par": rel(up1, fix["upar"][0]), "SijN_hyd": rel(s1, fix["shyd"][0]),
                  "yeta_mu1": abs(y1[0] - fix["yeta"][0][0]) / abs(fix["yeta"][0][0]),
                  "reference_O0_vs_O2_spread": {k: float(fix["spread_" + k][0])
                                                for k in ("conf", "upar", "shyd", "yeta")}}
        log(f"{label}: first step vs the reference: {parity}")
    for _ in range(args.warmup - 1):
        checked_step(T)
        T += args.dt
    kernels, dom_name, nbreak = {}, None, 2
    if not args.no_profile:
        rsd_amd.profile_enable(True)
        for _ in range(nbreak):
            checked_step(T)
            T += args.dt
        for name, (n, ms, fl, by) in rsd_amd.profile_read().items():
            if n:
                kernels[name] = {"launches_per_step": n / nbreak, "ms_per_step": ms / nbreak,
                                 "avg_launch_us": 1e3 * ms / n,
                                 "tflops": (fl / (ms * 1e-3) / 1e12) if fl else None,
                                 "GBps": (by / (ms * 1e-3) / 1e9) if by else None}
        dom_name = dom_family or max(kernels, key=lambda k: kernels[k]["ms_per_step"])
        rsd_amd.profile_enable(True, families=[dom_name])
    stage = np.zeros(6)
    nts, cg_iters, cg_fallbacks = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y, info = checked_step(T)
        stage += sus.stage_times()
        nts.append(sus.last_nt())
        it, fb, rd = sus.step_stats()
        cg_iters.append(it)
        cg_fallbacks.append(fb)
        redos += rd
        it_stats.append(sus.iterative_stats())
        T += args.dt
    elapsed = time.perf_counter() - t0
    log(f"{label}: timed {args.steps} steps: {1e3 * elapsed / args.steps:.2f} ms/step")
    prof = rsd_amd.profile_read() if not args.no_profile else {}
    rsd_amd.profile_enable(False)
    roof = None
    if dom_name and prof.get(dom_name, (0,))[0]:
        name, (n, ms, fl, by) = dom_name, prof[dom_name]
        if name.startswith("gemv_split") or name == "inner_gmres_graph":
            achieved = by / (ms * 1e-3)
            roof = {"kernel": name, "bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                    "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": measured_traffic(name),
                    "avg_launch_us": 1e3 * ms / n, "bytes_per_launch": by / n, "launches_per_step": n / args.steps}
        else:
            achieved = fl / (ms * 1e-3)
            roof = {"kernel": name, "bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                    "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": measured_traffic(name),
                    "avg_launch_us": 1e3 * ms / n, "flop_per_launch": fl / n, "launches_per_step": n / args.steps}
    stage /= args.steps
    return {"value": args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps, "parity": parity,
            "bad_info": bad_info, "redos": redos, "nts": nts, "cg_iters": cg_iters, "cg_fallbacks": cg_fallbacks,
            "it_stats": it_stats, "kernels": kernels, "nbreak": nbreak, "roofline": roof, "stage": stage,
            "T": T, "y": y}


def batch_line(rsd_amd, np, args):
    """Config 3: 64 independent bidisperse suspensions (N uniform in [100, 1000],
    phi in [0.1, 0.5], tests/golden/config3_batch.npz) stepped as one batch on
    the FP64 route; the width (concurrent host threads and streams) chosen by
    the batch autotuner; suspension-steps per second summed over the batch."""
    path = os.path.join(ROOT, "tests", "golden", "config3_batch.npz")
    if not os.path.exists(path):
        return None
    F = np.load(path)
    count = sum(1 for k in F.files if k.startswith("radii"))
    sus = [rsd_amd.Suspension(F[f"conf{k}"], F[f"radii{k}"], rsd_amd.default_config(float(F[f"L_phi_seed{k}"][0])))
           for k in range(count)]
    ns = [len(F[f"radii{k}"]) for k in range(count)]
    batch = rsd_amd.Batch(sus)
    best, ms = batch.autotune(args.dt, 0.0, widths=(1, 2, 4, 8, 16, 32))
    T, steps, bad = 0.0, args.batch_steps, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        _, info = batch.step(args.dt, T)
        bad += int(np.sum(info != 0))
        T += args.dt
    el = time.perf_counter() - t0
    batch.close()
    for x in sus:
        x.close()
    log(f"batch of {count}: width {best}, {1e3 * el / steps:.1f} ms a batch step")
    return {"suspensions": count, "particles": int(sum(ns)), "N_range": [min(ns), max(ns)],
            "value": count * steps / el, "unit": "suspension-steps/s", "ms_per_batch_step": 1e3 * el / steps,
            "width": best, "autotune_ms_by_width": ms,
            "sequential_ms_per_batch_step": ms.get(1), "steps": steps, "info_nonzero": bad,
            "route": "FP64 direct (Schur, CG sub-steps)"}


def learned_line(rsd_amd, np, args, gnn):
    """Config 4's evaluation at N=1000: GMRES iterations and solve times to a
    1e-10 relative residual on held-out suspensions (tests/golden/config4_suspensions.npz)."""
    path = os.path.join(ROOT, "tests", "golden", "config4_suspensions.npz")
    if gnn is None or not os.path.exists(path):
        return None
    F = np.load(path)
    names = {1: "block_jacobi", 3: "learned", 2: "mixed", 4: "mixed_learned"}
    rows = []
    for k in range(3):
        L, phi, seed = F[f"L_phi_seed{k}"]
        s = rsd_amd.Suspension(F[f"conf{k}"], F[f"radii{k}"], rsd_amd.default_config(float(L)))
        s.set_gnn(gnn)
        s.prepare_iterative(0.0)
        row = {"phi": float(phi), "N": 1000}
        for pc, nm in names.items():
            s.iterative_solve(pc, tol=1e-10, max_iter=2000)
            it, rr, ms = s.iterative_solve(pc, tol=1e-10, max_iter=2000)
            row[nm] = {"iterations": it, "ms": ms, "converged": rr <= 1e-10}
        rows.append(row)
        s.close()
        log(f"learned preconditioner at N=1000 phi={phi:.2f}: " +
            ", ".join(f"{nm} {row[nm]['iterations']} it {row[nm]['ms']:.1f} ms" for nm in names.values()))
    return rows


def config5_line(rsd_amd, np, args, gnn):
    """Config 5: N=4096 bidisperse at phi=0.5, two steps of dt=2e-4 on the
    direct FP64 route and on the mixed-precision learned iterative route,
    with the two routes' relative differences."""
    path = os.path.join(ROOT, "tests", "golden", "config5_suspension.npz")
    if gnn is None or not os.path.exists(path):
        return None
    F = np.load(path)
    L, dt = F["L_dt_phi_seed"][:2]
    out, last = {"N": 4096, "phi": 0.5, "system": 45056}, {}
    for route in ("fp64_direct", "mixed_learned"):
        s = rsd_amd.Suspension(F["conf0"], F["radii"], rsd_amd.default_config(float(L)))
        if route == "mixed_learned":
            s.set_resistance_route(s.RESIST_ITERATIVE)
            s.set_iterative(precond=s.PRECOND_MIXED_LEARNED, tol=args.gmres_tol, max_iter=2000)
            s.set_gnn(gnn)
        T, times, its = 0.0, [], []
        for _ in range(2):
            t0 = time.perf_counter()
            y, info = s.step(float(dt), T)
            times.append(time.perf_counter() - t0)
            fb = False
            if route == "mixed_learned":
                it, rr, fb = s.iterative_stats()
                its.append(it)
            if info != 0 or fb:
                s.close()
                return {"error": f"{route}: INFO {info}, fallback {fb}"}
            T += float(dt)
        up, sh, _ = s.last()
        last[route] = (up.copy(), sh.copy(), np.array(y[:3]))
        s.close()
        out[route] = {"s_per_step": times, "steps_per_s": 1.0 / times[-1]}
        if its:
            out[route]["gmres_outer_iterations"] = its
        log(f"config 5 (N=4096) {route}: {times[0]:.2f} s, {times[1]:.2f} s a step")
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(a))
    (u1, s1, y1), (u2, s2, y2) = last["fp64_direct"], last["mixed_learned"]
    out["relative_difference_after_two_steps"] = {"U_par": rel(u1, u2), "SijN_hyd": rel(s1, s2),
                                                  "yeta": rel(y1, y2)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the step runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding it over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied 1 GiB three times (reads 1048576 KiB and writes 1048576 KiB per dispatch)")
        return
    import numpy as np
    import rsd_amd
    from rsd_amd.configs import random_suspension

    rsd_amd.init(0)
    fix_path = os.path.join(ROOT, "tests", "golden", "config2_step.npz")
    fix = None
    if (args.n, args.phi, args.seed, args.dt) == (512, 0.4, 2000, 5e-4) and os.path.exists(fix_path):
        fix = np.load(fix_path)
        conf, radii = fix["conf0"].copy(), fix["radii"].copy()
        L = float(fix["L_dt_phi_seed"][0])
    else:
        conf, radii, L = random_suspension(args.n, args.phi, args.seed)
    log(f"suspension ready: N={args.n} phi={args.phi} L={L:.3f}")
    cfg = rsd_amd.default_config(L)

    def fail(res, label):
        log(f"{label}: steps with INFO != 0: {res['bad_info']}")
        print(json.dumps({"metric": METRIC, "value": None, "error": f"{label}: INFO != 0 on {len(res['bad_info'])} steps"}))
        sys.exit(3)

    # the plain FP64 path: Schur route, CG sub-steps
    sus = rsd_amd.Suspension(conf, radii, cfg)
    mode = {"direct": sus.SUBSTEP_DIRECT, "pcg": sus.SUBSTEP_PCG, "pcg-one-launch": sus.SUBSTEP_PCG_ONE_LAUNCH}
    if args.cg_tol is None:
        sus.set_substep_solver(mode[args.substep_solver])
    else:
        sus.set_substep_solver(mode[args.substep_solver], args.cg_tol)
    r64 = run_path(rsd_amd, np, sus, args, fix, "fp64")
    if r64["bad_info"]:
        fail(r64, "fp64")
    # the mixed-precision, learned-preconditioned iterative route
    rml, gnn, precond = None, None, None
    if os.path.exists(args.gnn_params):
        gnn = rsd_amd.Gnn.load(args.gnn_params)
    if not args.no_mixed:
        sml = rsd_amd.Suspension(conf, radii, cfg)
        sml.set_resistance_route(sml.RESIST_ITERATIVE)
        precond = sml.PRECOND_MIXED_LEARNED if gnn is not None else sml.PRECOND_MIXED
        sml.set_iterative(precond=precond, tol=args.gmres_tol, max_iter=400, inner_iters=args.inner_iters)
        if gnn is not None:
            sml.set_gnn(gnn)
        if args.no_graphs:
            sml.set_iterative_graphs(False)
        rml = run_path(rsd_amd, np, sml, args, fix, "mixed+learned",
                       dom_family="gemv_split_f32" if args.no_graphs else "inner_gmres_graph")
        sml.close()
        if rml["bad_info"]:
            fail(rml, "mixed+learned")
        fbs = sum(int(s[2]) for s in rml["it_stats"])
    # Ewald pair-image evaluations per second, parity mode (the reference's
    # rule: the lattice at the last timed step) and config mode (4 real images,
    # reciprocal vectors |n| <= 6)
    T = r64["T"]

    def pair_images(nr, ni, k2):
        nk = sum(1 for a in range(-ni, ni + 1) for b in range(-ni, ni + 1) for c in range(-ni, ni + 1)
                 if 0 < a * a + b * b + c * c and (k2 == 0 or a * a + b * b + c * c <= k2))
        return args.n * (args.n - 1) / 2 * ((2 * nr + 1) ** 3 + nk), nk
    lat = rsd_amd.set_lattice_skew(T - args.dt, np.full(3, L), 1.0)
    rsd_amd.calc_lattice_inverse(lat, 1e-4, radii)
    ewald = {}
    ev_p, _ = pair_images(lat.nr, lat.ni, 0)
    kernels = r64["kernels"]
    if kernels.get("ewald_pair_kernel"):
        ewald["parity"] = {"NR": lat.nr, "NI": lat.ni, "pair_image_evals_per_step": ev_p,
                           "ewald_pair_ms": kernels["ewald_pair_kernel"]["ms_per_step"],
                           "pair_image_evals_per_s": ev_p / (1e-3 * kernels["ewald_pair_kernel"]["ms_per_step"])}
    if args.config_mode_steps > 0:
        bad = []
        sus.set_ewald_images(4, 6, 36)
        rsd_amd.profile_enable(True, families=["ewald_pair_kernel"])
        t0 = time.perf_counter()
        for _ in range(args.config_mode_steps):
            _, info = sus.step(args.dt, T)
            if info != 0:
                bad.append(info)
            T += args.dt
        cm_el = time.perf_counter() - t0
        cprof = rsd_amd.profile_read()
        rsd_amd.profile_enable(False)
        sus.set_ewald_images()
        ev_c, nk = pair_images(4, 6, 36)
        n_l, ms_l = cprof["ewald_pair_kernel"][0], cprof["ewald_pair_kernel"][1]
        ewald["config"] = {"NR": 4, "NI": 6, "kmax2": 36, "k_vectors": nk, "steps": args.config_mode_steps,
                           "ms_per_step": 1e3 * cm_el / args.config_mode_steps,
                           "steps_per_s": args.config_mode_steps / cm_el, "pair_image_evals_per_step": ev_c,
                           "ewald_pair_ms": ms_l / max(n_l, 1),
                           "pair_image_evals_per_s": ev_c / (1e-3 * ms_l / max(n_l, 1))}
        log(f"config-mode Ewald (4 real images, |n| <= 6): {ewald['config']['ms_per_step']:.2f} ms/step")
        if bad:
            r64["bad_info"] = bad
            fail(r64, "fp64 config-mode")
    if not args.no_cpu_baseline:
        log("timing the reference on the host (bounded sample)")
    nt_mean = float(np.mean(r64["nts"]))
    base = None if args.no_cpu_baseline else cpu_baseline(conf, radii, L, nt_mean)
    stage = r64["stage"]
    steps_total = 1 + (args.warmup - 1) + (0 if args.no_profile else r64["nbreak"]) + args.steps
    fp64 = {"value": r64["value"], "unit": "SD steps/s", "ms_per_step": r64["ms_per_step"], "dtype": "f64",
            "path": "Schur-complement dense FP64 solve, CG sub-steps", "substep_solver": args.substep_solver,
            "nt_dem_mean": nt_mean, "cg_iterations_per_step": float(np.mean(r64["cg_iters"])),
            "cg_fallbacks": int(np.sum(r64["cg_fallbacks"])), "explicit_redo_steps": int(r64["redos"]),
            "stages_ms": {"ewald": stage[0], "inv11N": stage[1], "lub_R": stage[2], "inv6N": stage[3],
                          "matvec_update": stage[4], "step": stage[5]},
            "kernels": kernels, "roofline": r64["roofline"], "parity": r64["parity"],
            "stresslet_rel_err": r64["parity"]["SijN_hyd"] if r64["parity"] else None}
    mixed = None
    if rml is not None:
        its = [s[0] for s in rml["it_stats"]]
        mixed = {"value": rml["value"], "unit": "SD steps/s", "ms_per_step": rml["ms_per_step"],
                 "dtype": "f64 outer GMRES + f32 inner GMRES",
                 "path": "iterative route: FGMRES (FP64) on K x = [Fe; -EIN] with FP32 inner GMRES preconditioned by "
                         + ("the learned pair correction (" + os.path.relpath(args.gnn_params, ROOT) + ")"
                            if gnn is not None else "block-Jacobi (no trained model found)"),
                 "gmres_tol": args.gmres_tol, "inner_iters": args.inner_iters,
                 "gmres_outer_iterations_per_step": float(np.mean(its)),
                 "max_relres": float(max(s[1] for s in rml["it_stats"])), "direct_fallback_steps": fbs,
                 "nt_dem_mean": float(np.mean(rml["nts"])), "kernels": rml["kernels"], "roofline": rml["roofline"],
                 "parity": rml["parity"],
                 "stresslet_rel_err": rml["parity"]["SijN_hyd"] if rml["parity"] else None,
                 "speedup_vs_fp64_gpu": rml["value"] / r64["value"]}
    extra_batch = None if args.no_batch else batch_line(rsd_amd, np, args)
    extra_c4 = None if args.no_config4 else learned_line(rsd_amd, np, args, gnn)
    extra_c5 = None if args.no_config5 else config5_line(rsd_amd, np, args, gnn)
    head = mixed if (mixed is not None and args.headline == "mixed-learned") else None
    out = {
        "metric": METRIC, "value": head["value"] if head else fp64["value"], "unit": "SD steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"] if head else fp64["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": head["dtype"] if head else "f64", "data": "synthetic",
        "config": {"workload": f"N={args.n} monodisperse a=1, phi={args.phi}, cubic L={L:.3f}, shear 1, "
                               f"Lees-Edwards, Ewald LINERR=1e-4 (reference rule), lubrication tables, dt={args.dt}",
                   "N": args.n, "phi": args.phi, "seed": args.seed, "dt": args.dt,
                   "headline_path": "mixed+learned" if head else "fp64", "steps_total_per_path": steps_total},
        "roofline": head["roofline"] if head else fp64["roofline"], "cpu_baseline": base,
        "stresslet_rel_err": head["stresslet_rel_err"] if head else fp64["stresslet_rel_err"],
        "mixed_learned": mixed, "fp64_gpu": fp64,
        "speedup_mixed_learned_vs_fp64_gpu": mixed["speedup_vs_fp64_gpu"] if mixed else None,
        "ewald_modes": ewald, "batch_config3": extra_batch, "learned_config4": extra_c4,
        "config5": extra_c5, "last_yeta_mu": [float(v) for v in r64["y"][:2]],
    }
    print(json.dumps(out))
    sus.close()
    if gnn is not None:
        gnn.close()


if __name__ == "__main__":
    main()
