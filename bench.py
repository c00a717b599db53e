"""Benchmark: particle-steps/s of the coupled DEM-SPH dam break on one MI355X.

Workload: the coupled dam break at dx = 0.005 m — 512,000 fluid particles,
452,376 wall dummy particles and 2,000 DEM spheres (Hertz–Mindlin contact,
tank planes), synthetic and deterministic.  One step = the whole per-timestep
hot path (neighbour search, solid half-kick/drift, fluid integration with the
Riemann WCSPH pair loops, DEM contact forces, solid half-drift/kick), all
inputs resident in HBM.

Prints one JSON line.  Besides the headline (mixed precision) it reports the
all-FP64 GPU path on the same workload, the HBM roofline fraction of the
dominant kernel (from HIP-event kernel times taken inside the timed region)
and a CPU baseline (the FP64 OpenMP oracle on a bounded sample of steps).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dem-sphsource_amd"))

import numpy as np  # noqa: E402

METRIC = "particle-steps/s at 1M SPH+DEM; % HBM roofline; FP32-mixed vs FP64 speedup"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# Algorithmic bytes of the fluid pair loops per launch: every per-particle
# field the kernel uses read once, every output written once, plus the 4-byte
# neighbour index of each directed pair (neighbour fields served from cache
# are not counted).  Field bytes per fluid particle:
#   continuity: read pos+h (32, f64x4), v+rho0 (16 f32x4 | 32 f64x4),
#               dRho (4 f32 | 8 f64), c,nu,m,h (16 f32x4 | c,nu 16 f64x2),
#               list count (4) and fluid-solid offset (4);
#               write dRho f64 (8) [+ f32 shadow (4)]
#   momentum:   the same reads; write acceleration f64x4 (32) and the
#               fluid->solid force f64x4 (32) per fluid-solid pair
PAIR_BYTES = {
    "continuity": {"mixed": (32 + 16 + 4 + 16 + 4 + 4, 8 + 4, 0),
                   "fp64": (32 + 32 + 8 + 16 + 4 + 4, 8, 0)},
    "momentum": {"mixed": (32 + 16 + 4 + 16 + 4 + 4, 32, 32),
                 "fp64": (32 + 32 + 8 + 16 + 4 + 4, 32, 32)},
}


def build_scene(dx):
    import scenes
    return scenes.coupled_dam_break(dx=dx)


def run_gpu(scene, precision, steps, warmup, timing_steps, eager=False, flags=0, tuning=None):
    """Times `steps` steps after `warmup` (HIP-graph replay unless eager), then
    runs `timing_steps` more steps eagerly with per-family HIP events on the
    context stream (events cannot live inside a captured graph; kernel
    durations do not depend on how the kernels were launched).  `scene` may
    be a list of scenes: they are uploaded as one batch and step with the
    smallest of their time steps."""
    import dsph
    g = dsph.Solver(precision)
    g.set_flags(flags)
    if isinstance(scene, list):
        g.upload_batch(scene)
        dt = min(sc.dt for sc in scene)
    else:
        g.upload(scene)
        dt = scene.dt
    if tuning:
        g.apply_tuning(tuning)
    g.initialize()
    i = 1
    if warmup:
        g.run(warmup, dt, first=i, graphs=not eager)
        i += warmup
    g.synchronize()
    rb0 = int(g.status()[8])
    t0 = time.perf_counter()
    # with graphs the run ends with a status read (device sync)
    g.run(steps, dt, first=i, graphs=not eager)
    g.synchronize()
    t1 = time.perf_counter()
    g.rebuilds_timed = int(g.status()[8]) - rb0
    i += steps
    times = {}
    if timing_steps:
        g.set_timing(True)
        g.run(timing_steps, dt, first=i, graphs=False)
        g.synchronize()
        times = g.kernel_times()
        g.set_timing(False)
    return g, t1 - t0, times


def pair_counts(g):
    import dsph
    a, _, _ = g.list(dsph.LIST_FLUID_FLUID)
    fs, _, _ = g.list(dsph.LIST_FLUID_SOLID)
    return 2 * len(a), len(fs)


def roofline(times, prec_key, nf, ff_directed, fs_pairs):
    best = None
    for fam in ("continuity", "momentum"):
        if fam not in times or times[fam][1] == 0:
            continue
        tot_ms, launches = times[fam]
        rd, wr, per_fs = PAIR_BYTES[fam][prec_key]
        nbytes = nf * (rd + wr) + 4 * (ff_directed + fs_pairs) + per_fs * fs_pairs
        avg_s = tot_ms / launches / 1e3
        gbs = nbytes / avg_s / 1e9
        cand = {"kernel": fam, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "bytes_per_launch": int(nbytes), "avg_launch_ms": round(avg_s * 1e3, 4),
                "launches": launches, "total_ms": round(tot_ms, 3)}
        if best is None or tot_ms > best["total_ms"]:
            best = cand
    return best


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_pair_loops_coupled_dam_break.txt")


def traffic_from_profile(kernel, prec_key):
    """HBM bytes per launch of a kernel family from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy);
    lines read "TRAFFIC <family> <precision> <bytes>"."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == "TRAFFIC" and w[1] == kernel and w[2] == prec_key:
                    return int(float(w[3]))
    except OSError:
        pass
    return None


def copy_probe(mib):
    """A plain device copy of known size (read mib MiB, write mib MiB), the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "GBps": round(8 * n / dt / 1e9, 1)}), flush=True)


def cpu_baseline(scene, budget_s, max_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.Oracle(threads=threads)
    o.upload(scene)
    o.initialize()
    n = 0
    t0 = time.perf_counter()
    while n < max_steps:
        o.step(n + 1, scene.dt)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    o.close()
    return {"value": round(scene.n_particles * n / dt, 1), "unit": "particle-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"FP64 OpenMP oracle, first {n} steps of the same workload after its "
                      f"initial search ({dt:.1f} s)"}


def tune(work, precision, flags, args):
    """The autotuner's choice for one precision (None when disabled; the
    --tuning configuration when given)."""
    import dsph
    if args.tuning:
        cfg = json.loads(args.tuning)
        return cfg.get("fp64" if precision == dsph.PREC_FP64 else "mixed", cfg)
    if args.no_autotune:
        return None
    import dsph
    cfg, trials = dsph.autotune(work, precision, flags, trial_steps=30, warmup=5)
    print(f"# autotuned ({'mixed' if precision == dsph.PREC_MIXED else 'fp64'}): {cfg}",
          file=sys.stderr, flush=True)
    return cfg


def other_workload(args):
    """The batched ensemble or the granular column: one JSON line of the same
    shape (no CPU baseline: the oracle runs one scene at a time)."""
    import dsph
    import scenes
    if args.workload == "ensemble":
        scs = scenes.ensemble(args.scenes)
        flags = dsph.FLAG_ART_VISCOSITY
        nf = sum(sc.n_fluid for sc in scs)
        desc = (f"batched dam-break ensemble: {len(scs)} independent scenes in one context "
                f"(segmented grids), {nf} fluid + {sum(sc.n_solid for sc in scs)} wall/DEM, "
                f"fluid per scene log-uniform 2k-60k, DEM 0-500, tank 0.8-1.6 m, alpha 0.01-0.1")
        work = scs
    else:
        sc = scenes.granular_column(n_sphere=110000, n_clump=12500, n_bonded=40000)
        flags, nf = 0, 0
        desc = (f"granular column collapse: {sc.n_solid} spheres (110k free, 12.5k tetrahedral "
                f"clumps of 4, a 40k bonded block), no fluid")
        work, scs = sc, [sc]
    npart = sum(sc.n_particles for sc in scs)
    print(f"# workload: {desc}", file=sys.stderr, flush=True)
    tsteps = 0 if args.no_timing else args.timing_steps
    tun = tune(work, dsph.PREC_MIXED, flags, args)
    g, secs, times = run_gpu(work, dsph.PREC_MIXED, args.steps, args.warmup, tsteps, args.eager, flags,
                             tun)
    value = npart * args.steps / secs
    print(f"# mixed: {secs / args.steps * 1e3:.3f} ms/step  {value:.4g} particle-steps/s",
          file=sys.stderr, flush=True)
    g.close()
    fp64 = None
    if not args.no_fp64 and args.fp64_steps > 0:
        tun64 = tune(work, dsph.PREC_FP64, flags, args)
        g64, s64, _ = run_gpu(work, dsph.PREC_FP64, args.fp64_steps, args.warmup, 0, args.eager, flags,
                              tun64)
        fp64 = {"value": round(npart * args.fp64_steps / s64, 1),
                "ms_per_step": round(s64 / args.fp64_steps * 1e3, 4), "steps": args.fp64_steps,
                "tuning": tun64}
        g64.close()
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic: seeded SplitMix64 scenes",
        "config": {"workload": desc, "particles": npart,
                   "precision": "mixed (f32 pair loops, f64 integration)", "tuning": tun},
        "fp64_path": fp64,
        "mixed_over_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "kernel_ms": {k: [round(v[0] / max(v[1], 1), 4), v[1]] for k, v in times.items()},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--fp64-steps", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-max-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=100,
                    help="eager steps after the timed region that carry the per-kernel HIP events")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step's kernels directly instead of replaying HIP graphs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--workload", default="coupled", choices=["coupled", "ensemble", "granular"],
                    help="coupled: the 1M-particle coupled dam break (the headline); ensemble: "
                         "the batched dam-break ensemble; granular: the clump/bond column")
    ap.add_argument("--scenes", type=int, default=256, help="scenes of the ensemble workload")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes and skin instead of autotuning them")
    ap.add_argument("--tuning", default=None, metavar="JSON",
                    help='fixed configuration instead of autotuning, e.g. \'{"mixed": {"skin": 0.2, '
                         '"pair_block": 256, "build_block": 32}, "fp64": {...}}\' (profiling runs)')
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs one GPU; the multi-GPU (slab-decomposed) path is a "
              "design only — run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload != "coupled":
        other_workload(args)
        return

    import dsph
    scene = build_scene(args.dx)
    npart = scene.n_particles
    moving = scene.n_fluid + int(np.count_nonzero(np.asarray(scene.solid_inv_mass) > 0))
    print(f"# workload: {scene.n_fluid} fluid + {scene.n_solid} solid, dt={scene.dt:.3e}",
          file=sys.stderr, flush=True)

    tsteps = 0 if args.no_timing else args.timing_steps
    tun = tune(scene, dsph.PREC_MIXED, 0, args)
    g, secs, times = run_gpu(scene, dsph.PREC_MIXED, args.steps, args.warmup, tsteps, args.eager,
                             tuning=tun)
    value = npart * args.steps / secs
    print(f"# mixed: {secs / args.steps * 1e3:.3f} ms/step  {value:.4g} particle-steps/s",
          file=sys.stderr, flush=True)
    ff_directed, fs_pairs = pair_counts(g)
    g_rebuilds = g.rebuilds_timed
    roof = roofline(times, "mixed", scene.n_fluid, ff_directed, fs_pairs) if times else None
    if roof:
        roof["traffic"] = traffic_from_profile(roof["kernel"], "mixed")
    g.close()

    fp64 = None
    if not args.no_fp64 and args.fp64_steps > 0:
        tun64 = tune(scene, dsph.PREC_FP64, 0, args)
        g64, s64, t64 = run_gpu(scene, dsph.PREC_FP64, args.fp64_steps, args.warmup, tsteps,
                                args.eager, tuning=tun64)
        fp64 = {"value": round(npart * args.fp64_steps / s64, 1),
                "ms_per_step": round(s64 / args.fp64_steps * 1e3, 4), "steps": args.fp64_steps,
                "tuning": tun64, "superset_rebuilds": g64.rebuilds_timed}
        r64 = roofline(t64, "fp64", scene.n_fluid, ff_directed, fs_pairs) if t64 else None
        if r64:
            fp64["roofline_frac"] = r64["frac"]
            fp64["dominant_kernel_ms"] = r64["avg_launch_ms"]
        g64.close()
        print(f"# fp64: {fp64['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, args.cpu_seconds, args.cpu_max_steps)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: deterministic lattice dam break + seeded-jitter DEM lattice",
        "config": {"workload": f"coupled DEM-SPH dam break dx={args.dx}: {scene.n_fluid} fluid + "
                               f"{scene.n_solid - 2000} wall + 2000 DEM spheres",
                   "particles": npart, "precision": "mixed (f32 pair loops, f64 integration)",
                   "tuning": tun},
        "moving_particle_steps_per_s": round(moving * args.steps / secs, 1),
        "superset_rebuilds": g_rebuilds,
        "fp64_path": fp64,
        "mixed_over_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "roofline": roof,
        "kernel_ms": {k: [round(v[0] / max(v[1], 1), 4), v[1]] for k, v in times.items()},
        "pairs": {"fluid_fluid_directed": ff_directed, "fluid_solid": fs_pairs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
