"""Benchmark of the geodesic force path: MD steps/s and exact geodesic pairs/s on one MI355X.

Workload (default): bumpy sphere L6 (81,920 faces), N = 20,000, truncated Gaussian (s = 0.05, cutoff 0.15),
Nose-Hoover chain NVT at T = 0.5, dt = 0.005; one step = 2 walks + 1 neighbour search + 1 batched exact-geodesic force
evaluation, all resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1|2|3|5] [--no-cpu-baseline]
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, face, bary, vel, threads, budget_s=30.0, accuracy=True):
    """The FP64 CPU restatement of the reference path (oracle, reference Release flags), timed on this host with
    `threads` OpenMP threads, and on one thread (the reference's own execution model) over a smaller sample. With
    `accuracy`, the same restatement (IEEE build) also checks the GPU path from this state (accuracy_check)."""
    r = _cpu_time(w, face, bary, vel, threads, budget_s)
    if threads > 1:
        one = _cpu_time(w, face, bary, vel, 1, budget_s / 3)
        r["single_thread"] = {"value": one["value"], "unit": one["unit"], "cores": 1, "sample": one["sample"]}
    if accuracy:
        log("accuracy of the GPU path against the restatement")
        r["accuracy_of_gpu_path"] = accuracy_check(w, face, bary, vel, threads)
    return r


def replicas_cpu_baseline(w, face, bary, vel, threads, budget_s=30.0):
    """CPU baseline of a particle-group workload: each replica is one reference run on its own mesh, so one step of
    all replicas costs the sum of the replicas' step times, each timed like a single system (_cpu_time) with an
    equal share of the budget; at `threads` threads and, on a third of the budget, at one thread."""
    reps = w.meta["replicas"]

    def total(th, budget):
        t, n_src, n_all = 0.0, 0, 0
        for r, rep in enumerate(reps):
            p0, n = rep["particle_offset"], rep["n"]
            rw = dataclasses.replace(w, X=rep["X"], T=rep["T"], potential=w.meta["pots"][r], range=rep["sigma"],
                                     maximum_distance=w.meta["maxd"][r], rmax=rep["sigma"], meta={})
            one = _cpu_time(rw, face[p0:p0 + n] - rep["face_offset"], bary[p0:p0 + n], vel[p0:p0 + n], th, budget / len(reps))
            t += 1.0 / one["value"]
            n_src += one["sources_timed"]
            n_all += n
        return t, n_src, n_all

    t, n_src, n_all = total(threads, budget_s)
    r = dict(value=1.0 / t, unit="steps/s", cores=threads, kind="port",
             sample=(f"sum over the {len(reps)} replicas of one step each (each replica one system on its own mesh): force "
                     f"evaluation of {n_src} of {n_all} sources (reference semantics) scaled per replica, plus one walk "
                     f"pass; oracle built -O3 -march=x86-64-v3 -ffast-math; {threads} OpenMP threads"))
    if threads > 1:
        t1, n1, _ = total(1, budget_s / 3)
        r["single_thread"] = {"value": 1.0 / t1, "unit": "steps/s", "cores": 1,
                              "sample": f"as above on one thread, {n1} of {n_all} sources"}
    return r


def _cpu_time(w, face, bary, vel, threads, budget_s):
    from oracle import pyoracle

    om = pyoracle.OracleMesh(w.X, w.T, fast=True)
    P = (w.X[w.T[face]] * bary[:, :, None]).sum(1) / bary.sum(1)[:, None]
    lo, hi = np.minimum(w.X.min(0), 0), np.maximum(w.X.max(0), 0)
    if w.cell_list:
        off, idx, md = pyoracle.cell_candidates(P, lo, hi, w.range, w.range)
    else:
        N = len(face)
        off = np.arange(0, N * (N - 1) + 1, N - 1, dtype=np.int32)
        idx = np.array([j for i in range(N) for j in range(N) if j != i], np.int32)
        md = np.full(N, 1e20)
    N = len(face)
    kind = int(w.potential[0])
    params = np.array(list(w.potential[1:]) + [0.0] * (3 - len(w.potential[1:])), float)
    # probe cost on a few sources, then choose a sample that fits the budget
    probe = min(N, max(threads * 4, 64))
    sub_off = np.append(off[: probe + 1], np.full(N - probe, off[probe])).astype(np.int32)
    t0 = time.perf_counter()
    om.compute_forces(face, bary, sub_off, idx, md, kind, params, w.submeshing, w.maximum_distance, threads=threads,
                      want_pairs=False)
    per_src = (time.perf_counter() - t0) / probe
    n_src = int(min(N, max(probe, budget_s / max(per_src, 1e-9))))
    sub_off = np.append(off[: n_src + 1], np.full(N - n_src, off[n_src])).astype(np.int32)
    t0 = time.perf_counter()
    om.compute_forces(face, bary, sub_off, idx, md, kind, params, w.submeshing, w.maximum_distance, threads=threads,
                      want_pairs=False)
    t_force = (time.perf_counter() - t0) * N / n_src
    disp = vel * (0.5 * w.dt)
    t0 = time.perf_counter()
    om.move(face, bary, disp, vel, threads=threads)
    t_move = time.perf_counter() - t0
    moves = 2 if w.integrator == "nvt" else 1
    t_step = t_force + moves * t_move
    return dict(value=1.0 / t_step, unit="steps/s", cores=threads, kind="port",
                sample=(f"force evaluation of {n_src} of {N} sources (reference semantics: whole patch, no truncation) "
                        f"scaled to {N}, plus {moves} full walk pass(es); oracle built -O3 -march=x86-64-v3 -ffast-math; "
                        f"{threads} OpenMP threads"),
                pairs_per_s=float(off[-1]) / t_step, sources_timed=n_src)


def profiled_traffic(config):
    """HBM bytes per launch of the geodesic kernel from the committed counter summary of this workload (FETCH_SIZE and
    WRITE_SIZE passes, scaled by the same counters over a copy of known size), or None."""
    if config != 3:
        return None
    path = os.path.join(ROOT, "profiles", "geodesicForceLds_bumpy_sphere_L6_N20000.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def profiled_counters(config):
    """SQ counter ratios of the geodesic kernel from the committed profile of this workload (lines 'sq_<name>: value'),
    or None."""
    if config != 3:
        return None
    path = os.path.join(ROOT, "profiles", "geodesicForceLds_bumpy_sphere_L6_N20000_sq.md")
    out = {}
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("sq_") and ":" in line:
                    k, v = line.split(":", 1)
                    out[k.strip()[3:]] = float(v)
    except (OSError, ValueError):
        return None
    return out or None


def screened_leg(ctx, w, sysm, steps, plain_steps_s, config):
    """The learned screen: trained on the GPU on exact pairs of MD snapshots of two other systems of this configuration
    (seeds 13378, 13379), calibrated on snapshots of two more (13380, 13381), then timed on this workload from its
    current state. The training systems live in a context of their own, so that this workload's mesh and particles stay
    in place."""
    import curvedspacesim_amd as csa
    from curvedspacesim_amd import screen, workloads

    t0 = time.perf_counter()
    tctx = csa.Context(0)
    md = {} if config == 3 else dict(train_md_steps=(0, 20, 40), holdout_md_steps=(25, 50))
    built = screen.build(tctx, config=config, train_seeds=(13378, 13379), holdout_seeds=(13380, 13381), steps=2000, batch=8192, **md)
    tctx.close()
    t_build = time.perf_counter() - t0
    log(f"screen built in {t_build:.1f} s: threshold {built['threshold']:.3f}, {built['train_pairs']} training pairs")
    # check on this state: forces of the screened computation, and whether any dropped pair lies within the cutoff
    sysm.record_pairs(True)
    F0, E0 = sysm.compute_forces()
    d = sysm.pair_distances()
    X = sysm.screen_features()
    rc = screen.pair_ranges(sysm, w)
    sysm.record_pairs(False)
    dropped = screen.dropped(ctx, built["params"], built["threshold"], X)
    sysm.set_screen(built["params"], built["threshold"])
    F1, E1 = sysm.compute_forces()
    dF = float(np.max(np.abs(F1 - F0)) / np.abs(F0).max())
    dE = float(abs(E1 - E0) / abs(E0))
    # two windows that reach a target at the same FP64 distance may leave either's FP32 start tangent, so forces of
    # two computations of one state agree to ~1e-9 of the largest force, energies exactly
    match = dF <= 1e-6 and dE <= 1e-12
    workloads.step(sysm, w, 1)
    sysm.sync()
    sysm.profile(True)
    sysm.kernel_times()  # reset
    t0 = time.perf_counter()
    workloads.step(sysm, w, steps)
    sysm.sync()
    el = time.perf_counter() - t0
    kt = sysm.kernel_times()
    sysm.profile(False)
    st = sysm.screen_stats()
    sysm.set_screen(None)
    v = steps / el
    screen_ms = kt["screen"][0] / steps
    # only the pairs of the guard band (separation at least 0.985 of the cutoff) are scored; the others are kept
    # unscored. MLP forward: 2 (16 x 64 + 64 x 64 + 64) = 10,368 flop per scored pair, on the BF16 matrix cores
    scored = int(np.sum(X[:, 0] <= screen.GUARD))
    mlp_tflops = scored * 10368 / (screen_ms * 1e-3) / 1e12 if screen_ms > 0 else 0.0
    return {"value": v, "unit": "steps/s", "speedup_vs_plain": v / plain_steps_s, "pairs_solved_per_step": st["kept"],
            "candidate_pairs_per_step": st["candidates"], "dropped_fraction": 1 - st["kept"] / max(st["candidates"], 1),
            "forces_match_plain": match, "max_force_difference_rel": dF, "energy_difference_rel": dE,
            "dropped_within_cutoff": int(np.sum(dropped & (d < rc))),
            "beyond_cutoff_fraction": float(np.mean(d >= rc)), "threshold": built["threshold"],
            "training_pairs": built["train_pairs"],
            "holdout_beyond_cutoff_fraction": built["holdout_beyond"], "holdout_dropped_fraction": built["holdout_drop"],
            "train_and_calibrate_s": t_build, "kernel_ms_per_step": {k: val[0] / steps for k, val in kt.items()},
            "screen_ms_per_step": screen_ms, "pairs_scored_per_step": scored,
            "mfma": {"achieved": mlp_tflops, "peak": 2500.0, "unit": "TFLOP/s", "frac": mlp_tflops / 2500.0,
                     "note": "MLP forward flops of the scored pairs over the whole screening launch (band test of every "
                             "candidate, features and MLP of the band's pairs); the matrix-core rate of the MLP itself is "
                             "that of the forward kernel over all candidates, in profiles/"}}


def accuracy_check(w, face, bary, vel, threads):
    """Errors against the oracle from the benchmark's current state, in a context of its own: candidate lists, pair
    distances (relative over pairs within the cutoff with d > 1e-3 rc, absolute below) and forces of one computation,
    then positions and velocities after one step taken by both from that state (fresh thermostat chains)."""
    import curvedspacesim_amd as csa
    from curvedspacesim_amd import workloads
    from oracle.pyoracle import OracleMesh, OracleSystem

    t0 = time.perf_counter()
    actx = csa.Context(0)
    ws = dataclasses.replace(w, face=face, bary=bary, vel=vel)
    sysm = workloads.make_system(actx, ws)
    sysm.record_pairs(True)
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    d = sysm.pair_distances()
    om = OracleMesh(w.X, w.T)
    osys = OracleSystem(om, face, bary, vel, w.potential, w.range, w.cell_list, w.submeshing, w.maximum_distance,
                        threads=threads)
    r = osys.compute_forces()
    same_lists = bool(np.array_equal(off, osys.last["off"]) and np.array_equal(idx, osys.last["idx"]))
    do = r["dist"]
    within = do < w.range
    big = within & (do > 1e-3 * w.range)
    rel = float(np.max(np.abs(d[big] - do[big]) / do[big])) if np.any(big) else 0.0
    small = within & ~big
    absd = float(np.max(np.abs(d[small] - do[small]))) if np.any(small) else 0.0
    fl = r["flags"]
    tie = np.array([np.any(fl[off[i]:off[i + 1]] & 2) for i in range(len(face))])  # cut-locus pairs (tangent ambiguous)
    scale = np.abs(r["forces"]).max()
    dF = np.abs(F - r["forces"]).max(1) / scale
    sysm.record_pairs(False)
    workloads.step(sysm, ws, 1)
    if w.integrator == "nvt":
        osys.nvt_step(w.dt, w.temperature, w.tau, w.chain)
    else:
        osys.nve_step(w.dt)
    f1, b1, v1, _ = sysm.get()
    dx = float(np.max(np.abs(om.positions(f1, b1) - om.positions(osys.face, osys.bary))))
    dv = float(np.max(np.abs(v1 - osys.vel)) / np.abs(osys.vel).max())
    actx.close()
    return {"pairs_compared": int(np.sum(within)), "candidate_lists_identical": same_lists,
            "max_rel_distance_error": rel, "max_abs_distance_error_small_d": absd,
            "max_force_error_rel": float(dF.max()), "max_force_error_rel_excluding_cut_locus": float(dF[~tie].max()),
            "sources_with_cut_locus_pairs": int(tie.sum()), "one_step_max_position_error": dx,
            "one_step_max_velocity_error_rel": dv, "oracle_threads": threads, "seconds": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-screen", action="store_true", help="skip the learned-screen leg")
    ap.add_argument("--no-graph", action="store_true", help="launch every kernel directly instead of replaying a captured step")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the CPU baseline's check of the GPU path")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="first copy 256 MiB between two device buffers (a kernel of known traffic, to scale the HBM "
                         "byte counters of a profile of this command)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}): this build measures one GPU's share of the "
              "work; sharding over several GPUs is designed but not built", file=sys.stderr)
        sys.exit(2)

    if args.calibration_copy:  # before libcsg.so loads, so that both use the HIP runtime torch brings
        import torch

        a = torch.ones(1 << 26, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 268,435,456 bytes read and written")
        del a, b

    import curvedspacesim_amd as csa
    from curvedspacesim_amd import workloads
    w = workloads.config(args.config)
    log(f"workload: {w.name}")
    ctx = csa.Context(0)
    sysm = workloads.make_system(ctx, w)
    chosen, tune_ms = sysm.autotune()
    log(f"autotune: LDS kernel variant {chosen}; ms per force computation {tune_ms}")
    sysm.use_graph(not args.no_graph)
    log(f"warmup {args.warmup} steps")
    workloads.step(sysm, w, args.warmup)
    sysm.sync()
    st0 = sysm.status()
    log(f"timing {args.steps} steps ({'one captured step replayed' if not args.no_graph else 'direct launches'})")
    t0 = time.perf_counter()
    workloads.step(sysm, w, args.steps)
    sysm.sync()
    elapsed = time.perf_counter() - t0
    # the same number of steps with HIP events around every launch group (direct launches): per-kernel durations
    log(f"timing {args.steps} steps with per-kernel events")
    sysm.profile(True)
    sysm.kernel_times()  # reset
    t1 = time.perf_counter()
    workloads.step(sysm, w, args.steps)
    sysm.sync()
    elapsed_eager = time.perf_counter() - t1
    kt = sysm.kernel_times()
    sysm.profile(False)
    st = sysm.status()
    ps = sysm.patch_stats()
    if st["cand_overflow"] or st["unsolved"] or st["walk_failures"]:
        raise SystemExit(f"step errors: {st}")
    pairs = st["pairs"]
    steps_s = args.steps / elapsed
    forces_per_step = 1
    # algorithmic bytes of one launch of the geodesic kernel: per source the patch's faces (3 corner ids + 3
    # neighbour ids, 24 B), its vertices (FP64 xyz, 24 B), its targets (index + face + FP64 barycentrics, 32 B), the
    # source's own state (28 B) and its force (24 B)
    if kt["geodesic_lds"][1] > 0:  # patches (also whole-mesh mode with a finite truncation radius)
        geo_key, geo_kernel = "geodesic_lds", "geodesicForceLds"
        bytes_launch = 24.0 * ps["faces"] + 24.0 * ps["verts"] + 32.0 * pairs + 52.0 * len(w.face)
    else:  # whole-mesh solves on HBM tables (512 lanes per source): every source reads the whole mesh
        geo_key, geo_kernel = "geodesic_hbm", "geodesicForceScratch<512> (whole mesh)"
        bytes_launch = len(w.face) * (24.0 * len(w.T) + 24.0 * len(w.X)) + 32.0 * pairs + 52.0 * len(w.face)
    geo_ms, geo_n = kt[geo_key]
    avg_ms = geo_ms / max(geo_n, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    out = {
        "metric": "MD steps/s + exact geodesic pairs/s at N=20k; screened vs plain, same error",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 distances/forces, fp32 tangents",
        "data": "synthetic",
        "config": {"workload": w.name, "N": int(len(w.face)), "faces": int(len(w.T)), "cutoff": w.range,
                   "integrator": w.integrator, "dt": w.dt, "screen": "plain (no learned screen)",
                   "lds_kernel_variant": csa.LDS_VARIANT_NAMES[chosen]},
        "autotune_ms_per_force": {n: float(t) for n, t in zip(csa.LDS_VARIANT_NAMES, tune_ms)},
        "graph_mode": not args.no_graph,
        "steps_per_s_direct_launches": args.steps / elapsed_eager,
        "pairs_per_step": int(pairs) * forces_per_step,
        "pairs_per_s": pairs * forces_per_step * steps_s,
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in kt.items()},
        "mean_patch_faces": ps["faces"] / len(w.face),
        "max_patch_faces": ps["max_faces"],
        "mean_fifo_steps": ps["steps"] / len(w.face),
        "mean_work_items": ps["items"] / len(w.face),
        "mean_source_us": {k: ps["us_" + k] / len(w.face) for k in ("patch", "propagate", "epilogue")},
        "lds_fallbacks_per_step": {k: v / (args.steps + args.warmup) for k, v in sysm.fallback_counts().items()},
        # the kernel is bound by the latency of its dependent LDS and memory round trips (the SQ counters of the committed
        # profile: waves waiting most of their cycles, VALU and FP64 issue far from peak); its HBM roofline is reported
        "roofline": {"kernel": geo_kernel, "bound": "latency", "peak_basis": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.config), "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": bytes_launch, "counters": profiled_counters(args.config)},
    }
    if not args.no_screen and args.config in (3, 5):
        out["screened"] = screened_leg(ctx, w, sysm, args.steps, steps_s, args.config)
    if not args.no_cpu_baseline and w.meta.get("groups") is not None:
        f, b, v, _ = sysm.get()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline with {threads} threads, replica by replica")
        out["cpu_baseline"] = replicas_cpu_baseline(w, f, b, v, threads)
        out["speedup_vs_cpu"] = steps_s / out["cpu_baseline"]["value"]
    elif not args.no_cpu_baseline:
        f, b, v, _ = sysm.get()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline with {threads} threads")
        out["cpu_baseline"] = cpu_baseline(w, f, b, v, threads, accuracy=not args.no_accuracy)
        out["speedup_vs_cpu"] = steps_s / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
