"""Benchmark: particle-steps/s of the coupled DEM-SPH dam break on one MI355X.

Workload: the coupled dam break at dx = 0.005 m — 512,000 fluid particles,
452,376 wall dummy particles and 2,000 DEM spheres (Hertz-Mindlin contact,
tank planes), synthetic and deterministic.  One step = the whole per-timestep
hot path (neighbour search, solid half-kick/drift, fluid integration with the
Riemann WCSPH pair loops, DEM contact forces, solid half-drift/kick), all
inputs resident in HBM.

Prints one JSON line.  Timed: --steps steps replayed from HIP graphs after
--warmup steps, for the mixed path (`value`) and, over the same step window,
for the all-FP64 path.  Then the same window runs again from a checkpoint
of its first step, eagerly, with HIP events around every kernel family on
the context stream: those durations give the dominant kernel's roofline.
A 1000-step window of both precisions (superset rebuilds amortised), a
CPU baseline (the FP64 OpenMP oracle on a bounded sample) and the VTU frames
leg (what writing frames costs the step loop, against the reference's
synchronous ASCII output of the same frame) are reported beside it.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dem-sphsource_amd"))

import numpy as np  # noqa: E402

METRIC = "particle-steps/s at 1M SPH+DEM; % HBM roofline; FP32-mixed vs FP64 speedup"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# Algorithmic bytes of the fluid pair loops per launch (the survey's
# measurement convention): every per-particle field of the reference's
# dataflow read once and every output written once, per fluid particle —
#   continuity (K8, per pass): position, velocity, h, rho0, dRho, c in and
#     dRho out: 88 B FP64, 44 B with the FP32 shadows of the mixed path;
#   momentum (K9): those plus nu in and the acceleration out: 112 B / 56 B;
# plus the 4-byte index of every directed fluid-fluid pair and every
# fluid-solid pair the pass visits, and for momentum the 24-byte force of
# every fluid-solid pair (calFluid2SolidForce's output, fused into it).
FIELD_BYTES = {"continuity": {"mixed": 44, "fp64": 88}, "momentum": {"mixed": 56, "fp64": 112}}
FS_FORCE_BYTES = 24


def build_scene(dx):
    import scenes
    return scenes.coupled_dam_break(dx=dx)


def new_solver(scene, precision, flags=0, tuning=None):
    import dsph
    g = dsph.Solver(precision)
    g.set_flags(flags)
    if isinstance(scene, list):
        g.upload_batch(scene)
        g.set_scene_time_steps([sc.dt for sc in scene])  # each scene at its own step
    else:
        g.upload(scene)
    if tuning:
        g.apply_tuning(tuning)
    g.initialize()
    return g


def timed_window(scene, precision, steps, warmup, timing=True, flags=0, tuning=None, graphs=True):
    """Graph-replayed steps warmup+1 .. warmup+steps on a fresh context:
    (seconds, superset rebuilds in the window, kernel-family times).  The
    kernel times come from the same window run again from a checkpoint of
    its first step, eagerly, with HIP events on the context stream."""
    import dsph
    dt = 0.0 if isinstance(scene, list) else scene.dt  # a batch steps each scene by its own
    g = new_solver(scene, precision, flags, tuning)
    if warmup:
        g.run(warmup, dt, first=1, graphs=graphs)
    blob = g.save_state() if timing else None
    g.synchronize()
    rb0 = int(g.status()[dsph.ST_REBUILDS])
    t0 = time.perf_counter()
    g.run(steps, dt, first=warmup + 1, graphs=graphs)
    g.synchronize()
    secs = time.perf_counter() - t0
    rebuilds = int(g.status()[dsph.ST_REBUILDS]) - rb0
    times = {}
    if timing:
        g.load_state(blob)
        g.set_timing(True)
        g.run(steps, dt, first=warmup + 1, graphs=False)
        g.synchronize()
        times = g.kernel_times()
        g.set_timing(False)
    return g, secs, rebuilds, times


def pair_counts(g):
    import dsph
    a, _, _ = g.list(dsph.LIST_FLUID_FLUID)
    fs, _, _ = g.list(dsph.LIST_FLUID_SOLID)
    return 2 * len(a), len(fs)


def candidates(g, ff_directed):
    """Candidates per fluid particle the pair loops evaluate (its group's
    superset row) against its neighbours in the fluid-fluid list."""
    st = g.superset_stats()
    n = max(st["particles"], 1)
    return {"walked_per_particle": round(st["walked"] / n, 2),
            "neighbours_per_particle": round(ff_directed / n, 2),
            "stored_per_particle": round(st["stored"] / n, 2), "longest_row": st["longest"]}


def roofline(times, prec_key, nf, ff_directed, fs_pairs):
    best = None
    for fam in ("continuity", "momentum"):
        if fam not in times or times[fam][1] == 0:
            continue
        tot_ms, launches = times[fam]
        nbytes = nf * FIELD_BYTES[fam][prec_key] + 4 * (ff_directed + fs_pairs)
        if fam == "momentum":
            nbytes += FS_FORCE_BYTES * fs_pairs
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, sample_steps, runs):
    """The FP64 OpenMP oracle on the box's host cores: one warm-up step after
    its initial search, then `runs` timed samples of `sample_steps` steps
    each; the median rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.Oracle(threads=threads)
    o.upload(scene)
    o.initialize()
    step = 1
    o.step(step, scene.dt)
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(sample_steps):
            step += 1
            o.step(step, scene.dt)
        rates.append(scene.n_particles * sample_steps / (time.perf_counter() - t0))
    o.close()
    med = statistics.median(rates)
    return {"value": round(med, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"FP64 OpenMP oracle on the same workload: 1 warm-up step after the initial "
                      f"search, then the median of {runs} runs of {sample_steps} steps "
                      f"(rates {[round(r) for r in rates]})"}


def _dir_bytes(d):
    return sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))


def frames_leg(scene, tuning, args):
    """VTU frames of the same workload, mixed precision.  The step loop
    (chunks of --frame-every graph-replayed steps) runs twice from one
    checkpoint: without frames, then with a frame of every species after
    each chunk (dsph_frame_begin, and dsph_frame_write on a writer thread
    per slot, as the C++ driver does): the difference is what a frame costs
    the loop.  Then one frame timed alone, and the reference's way of
    writing the same frame (solverBase.cpp:69-333: the FP64 state read back
    synchronously, then every value printed as ASCII, by the oracle's
    restatement of its writers) on the host."""
    import shutil
    import tempfile
    import threading

    import dsph
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    dt, every, nfr = scene.dt, args.frame_every, args.frames
    g = new_solver(scene, dsph.PREC_MIXED, tuning=tuning)
    g.run(args.warmup, dt, first=1, graphs=True)
    blob = g.save_state()
    first = args.warmup + 1

    def loop(frames_dir):
        # from the checkpoint (which drops the step graphs), one untimed
        # chunk records the graphs; then nfr timed chunks, each followed by
        # a frame when frames_dir is given: both loops do the same steps
        g.load_state(blob)
        g.run(every, dt, first=first, graphs=True)
        g.synchronize()
        writers, wtimes = [None, None], []

        def write(slot, k):
            t = time.perf_counter()
            g.frame_write(slot, frames_dir, k)
            wtimes.append(time.perf_counter() - t)

        t0 = time.perf_counter()
        for k in range(1, nfr + 1):
            g.run(every, dt, first=first + k * every, graphs=True)
            if frames_dir:
                slot = k % dsph.FRAME_SLOTS
                if writers[slot]:
                    writers[slot].join()
                step = first + (k + 1) * every - 1
                g.frame_begin(slot, step, step * dt)
                writers[slot] = threading.Thread(target=write, args=(slot, k))
                writers[slot].start()
        g.synchronize()
        t_loop = time.perf_counter() - t0
        for w in writers:
            if w:
                w.join()
        return t_loop, wtimes

    d = tempfile.mkdtemp(prefix="dsph_frames_")
    try:
        base, with_f, wtimes = [], [], []
        for _ in range(2):  # alternated, the faster of each kept
            base.append(loop(None)[0])
            tf, wt = loop(d)
            with_f.append(tf)
            wtimes += wt
        base, with_f = min(base), min(with_f)
        frame_bytes = _dir_bytes(d) / nfr
        shutil.rmtree(d)
        os.mkdir(d)
        # one frame alone: the begin call, then the write (copy wait, pair
        # copy, files); the packing kernels and the copy timed with HIP
        # events on their streams
        g.synchronize()
        g.set_timing(True)
        t0 = time.perf_counter()
        g.frame_begin(0, first, first * dt)
        t1 = time.perf_counter()
        g.frame_write(0, d, 0)
        t2 = time.perf_counter()
        ft = g.kernel_times()
        g.set_timing(False)
        ns_pairs = len(g.list(dsph.LIST_FLUID_SOLID)[0])
        d2h = 36 * scene.n_fluid + 60 * scene.n_solid + 32 * ns_pairs
        d2h_begin = 36 * scene.n_fluid + 60 * scene.n_solid  # copied beside the steps
        # the reference's way: FP64 read-back of what its writers print,
        # then the ASCII files
        t3 = time.perf_counter()
        fl, so = g.fluid(), g.solid()
        fi, sj, _ = g.list(dsph.LIST_FLUID_SOLID)
        F = g.fs_force()
        t4 = time.perf_counter()
        ns = scene.n_solid
        clus = scene.solid_cluster if scene.solid_cluster is not None else -np.ones(ns)
        clump = scene.solid_clump if scene.solid_clump is not None else -np.ones(ns)
        ref = os.path.join(d, "ref")
        os.mkdir(ref)
        oracle.vtu_fluid_ascii(os.path.join(ref, "fluid_0000.vtu"), first * dt, first, fl["pos"], scene.fluid_h,
                               fl["vel"], scene.fluid_rho0, fl["drho"], scene.fluid_c)
        oracle.vtu_solid_ascii(os.path.join(ref, "solid_0000.vtu"), first * dt, first, so["pos"],
                               scene.solid_radius, clus, clump, so["vel"], so["angvel"], so["normal"])
        oracle.vtu_fluid_solid_ascii(os.path.join(ref, "fluid2Solid_0000.vtu"), first * dt, first, fi, sj,
                                     fl["pos"], so["pos"], F)
        t5 = time.perf_counter()
        ascii_bytes = _dir_bytes(ref)
    finally:
        shutil.rmtree(d, ignore_errors=True)
        g.close()
    per_frame = (with_f - base) / nfr
    pack_ms, copy_ms = ft["frame_pack"][0], ft["frame_copy"][0]
    # algorithmic bytes of the packing kernels: per fluid particle 84 B read
    # (position+h, velocity+rho0, dRho, c, caller index) and 36 B written,
    # per solid 136 B read (position+radius, velocity, angular velocity,
    # normal, cluster, clump) and 60 B written, per pair 104 B read (two
    # indices, both positions, the force) and 32 B written
    pack_bytes = 120 * scene.n_fluid + 196 * scene.n_solid + 136 * ns_pairs
    out = {
        "every_steps": every, "frames": nfr, "species": "fluid, solid, fluid2Solid",
        "frame_bytes": int(frame_bytes), "d2h_bytes": d2h, "fluid_solid_pairs": ns_pairs,
        "loop_ms_per_step_without": round(base / (nfr * every) * 1e3, 4),
        "loop_ms_per_step_with": round(with_f / (nfr * every) * 1e3, 4),
        "loop_ms_per_frame": round(per_frame * 1e3, 3),
        "write_ms_on_thread": round(statistics.median(wtimes) * 1e3, 2),
        "alone": {"begin_ms": round((t1 - t0) * 1e3, 3), "write_ms": round((t2 - t1) * 1e3, 2),
                  "MB_per_s": round(frame_bytes / (t2 - t0) / 1e6, 1),
                  "pack_ms": round(pack_ms, 4), "pack_GB_per_s": round(pack_bytes / pack_ms / 1e6, 1),
                  "pack_bytes": pack_bytes, "copy_ms": round(copy_ms, 4),
                  "copy_GB_per_s": round(d2h_begin / copy_ms / 1e6, 1)},
        "reference_way": {"readback_ms": round((t4 - t3) * 1e3, 2), "ascii_write_ms": round((t5 - t4) * 1e3, 1),
                          "frame_bytes": int(ascii_bytes), "kind": "port",
                          "note": "FP64 read-back through the C ABI, then the oracle's restatement of "
                                  "the reference's ASCII writers on one host thread; the GPU waits"},
        "files": "written to a temporary directory and removed",
    }
    print(f"# frames: {out['loop_ms_per_frame']} ms of step loop per frame (every {every} steps), "
          f"write {out['write_ms_on_thread']} ms on the writer thread, {frame_bytes / 1e6:.1f} MB; "
          f"reference way {out['reference_way']['readback_ms'] + out['reference_way']['ascii_write_ms']:.0f} ms, "
          f"{ascii_bytes / 1e6:.1f} MB", file=sys.stderr, flush=True)
    return out


def tune(work, precision, flags, args):
    """The autotuner's choice for one precision (None when disabled; the
    --tuning configuration when given)."""
    import dsph
    if args.tuning:
        cfg = json.loads(args.tuning)
        return cfg.get("fp64" if precision == dsph.PREC_FP64 else "mixed", cfg)
    if args.no_autotune:
        return None
    cfg, trials = dsph.autotune(work, precision, flags, trial_steps=40, warmup=5,
                                advance=args.tune_advance)
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
    tun = tune(work, dsph.PREC_MIXED, flags, args)
    g, secs, rb, times = timed_window(work, dsph.PREC_MIXED, args.steps, args.warmup,
                                      not args.no_timing, flags, tun)
    g.close()
    value = npart * args.steps / secs
    print(f"# mixed: {secs / args.steps * 1e3:.3f} ms/step  {value:.4g} particle-steps/s",
          file=sys.stderr, flush=True)
    fp64 = None
    if not args.no_fp64:
        tun64 = tune(work, dsph.PREC_FP64, flags, args)
        g64, s64, rb64, _ = timed_window(work, dsph.PREC_FP64, args.steps, args.warmup, False, flags, tun64)
        g64.close()
        fp64 = {"value": round(npart * args.steps / s64, 1), "ms_per_step": round(s64 / args.steps * 1e3, 4),
                "steps": args.steps, "superset_rebuilds": rb64, "tuning": tun64}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic: seeded SplitMix64 scenes",
        "config": {"workload": desc, "particles": npart,
                   "precision": "mixed (f32 pair loops, f64 integration)", "tuning": tun},
        "superset_rebuilds": rb,
        "fp64_path": fp64,
        "mixed_over_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "kernel_ms": {k: [round(v[0] / max(v[1], 1), 4), v[1]] for k, v in times.items()},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def fp64_only(scene, args):
    """The all-FP64 leg alone (a profiler run keeps one context per process):
    a JSON line with its step time, kernel-family times and roofline."""
    import dsph
    tun64 = tune(scene, dsph.PREC_FP64, 0, args)
    g, secs, rb, times = timed_window(scene, dsph.PREC_FP64, args.steps, args.warmup, not args.no_timing,
                                      tuning=tun64, graphs=not args.eager)
    ff_directed, fs_pairs = pair_counts(g)
    g.close()
    roof = roofline(times, "fp64", scene.n_fluid, ff_directed, fs_pairs) if times else None
    if roof:
        roof["traffic"] = traffic_from_profile(roof["kernel"], "fp64")
    print(json.dumps({"fp64_only": True, "ms_per_step": round(secs / args.steps * 1e3, 4),
                      "steps": args.steps, "warmup": args.warmup, "superset_rebuilds": rb,
                      "tuning": tun64, "roofline": roof,
                      "kernel_ms": {k: [round(v[0] / max(v[1], 1), 4), v[1]] for k, v in times.items()}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--steady-steps", type=int, default=1000,
                    help="also time this many steps per precision (superset rebuilds amortised); 0: off")
    ap.add_argument("--cpu-sample-steps", type=int, default=4)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-frames", action="store_true", help="skip the VTU frames leg")
    ap.add_argument("--frames-only", action="store_true",
                    help="only the VTU frames leg, printed as its own JSON line (profiler runs)")
    ap.add_argument("--frame-every", type=int, default=100, help="steps between frames in the frames leg")
    ap.add_argument("--frames", type=int, default=4, help="frames in the frames leg")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-timing", action="store_true",
                    help="skip the eager re-run of the window with per-kernel HIP events")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step's kernels directly instead of replaying HIP graphs "
                         "(profiler counter runs)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="only the all-FP64 leg, one context per process (profiler runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--workload", default="coupled", choices=["coupled", "ensemble", "granular"],
                    help="coupled: the 1M-particle coupled dam break (the headline); ensemble: "
                         "the batched dam-break ensemble; granular: the clump/bond column")
    ap.add_argument("--scenes", type=int, default=256, help="scenes of the ensemble workload")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes and skin instead of autotuning them")
    ap.add_argument("--tune-advance", type=int, default=400,
                    help="steps the autotuner runs before its trials, so they time the moving flow")
    ap.add_argument("--tuning", default=None, metavar="JSON",
                    help='fixed configuration instead of autotuning, e.g. \'{"mixed": {"skin": 0.2, '
                         '"pair_block": 256, "build_block": 32, "group": 2}, "fp64": {...}}\'')
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

    if args.fp64_only:
        fp64_only(scene, args)
        return
    if args.frames_only:
        tn = tune(scene, dsph.PREC_MIXED, 0, args)
        print(json.dumps({"frames": frames_leg(scene, tn, args)}), flush=True)
        return

    tun = tune(scene, dsph.PREC_MIXED, 0, args)
    g, secs, rebuilds, times = timed_window(scene, dsph.PREC_MIXED, args.steps, args.warmup,
                                            not args.no_timing, tuning=tun, graphs=not args.eager)
    value = npart * args.steps / secs
    print(f"# mixed: {secs / args.steps * 1e3:.3f} ms/step  {value:.4g} particle-steps/s, "
          f"{rebuilds} rebuilds", file=sys.stderr, flush=True)
    ff_directed, fs_pairs = pair_counts(g)
    cand = candidates(g, ff_directed)
    tiles = g.tile_stats()
    roof = roofline(times, "mixed", scene.n_fluid, ff_directed, fs_pairs) if times else None
    if roof:
        roof["traffic"] = traffic_from_profile(roof["kernel"], "mixed")
    g.close()

    fp64 = None
    if not args.no_fp64:
        tun64 = tune(scene, dsph.PREC_FP64, 0, args)
        g64, s64, rb64, t64 = timed_window(scene, dsph.PREC_FP64, args.steps, args.warmup,
                                           not args.no_timing, tuning=tun64, graphs=not args.eager)
        g64.close()
        fp64 = {"value": round(npart * args.steps / s64, 1),
                "ms_per_step": round(s64 / args.steps * 1e3, 4), "steps": args.steps,
                "warmup": args.warmup, "superset_rebuilds": rb64, "tuning": tun64}
        r64 = roofline(t64, "fp64", scene.n_fluid, ff_directed, fs_pairs) if t64 else None
        if r64:
            fp64["roofline_frac"] = r64["frac"]
            fp64["dominant_kernel_ms"] = r64["avg_launch_ms"]
        print(f"# fp64: {fp64['ms_per_step']:.3f} ms/step, {rb64} rebuilds", file=sys.stderr, flush=True)

    steady = None
    if args.steady_steps > 0:
        steady = {"steps": args.steps if args.steps >= args.steady_steps else args.steady_steps,
                  "warmup": args.warmup}
        legs = [("mixed", dsph.PREC_MIXED, tun)]
        if fp64:
            legs.append(("fp64", dsph.PREC_FP64, fp64["tuning"]))
        for name, prec, tn in legs:
            if args.steps >= args.steady_steps:
                ms = secs / args.steps * 1e3 if name == "mixed" else fp64["ms_per_step"]
                rb = rebuilds if name == "mixed" else fp64["superset_rebuilds"]
            else:
                gs, ss, rb, _ = timed_window(scene, prec, args.steady_steps, args.warmup, False, tuning=tn)
                gs.close()
                ms = ss / args.steady_steps * 1e3
            steady[name] = {"ms_per_step": round(ms, 4), "value": round(npart / ms * 1e3, 1),
                            "superset_rebuilds": rb}
            print(f"# steady {name}: {ms:.3f} ms/step over {steady['steps']} steps, {rb} rebuilds",
                  file=sys.stderr, flush=True)
        if "fp64" in steady:
            steady["mixed_over_fp64"] = round(steady["fp64"]["ms_per_step"] / steady["mixed"]["ms_per_step"], 3)

    frames = None if args.no_frames else frames_leg(scene, tun, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, args.cpu_sample_steps, args.cpu_runs)

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
        "superset_rebuilds": rebuilds,
        "fp64_path": fp64,
        "mixed_over_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "steady_state": steady,
        "roofline": roof,
        "kernel_ms": {k: [round(v[0] / max(v[1], 1), 4), v[1]] for k, v in times.items()},
        "pairs": {"fluid_fluid_directed": ff_directed, "fluid_solid": fs_pairs},
        "candidates": cand,
        "tiles": tiles,
        "cpu_baseline": cpu,
        "frames": frames,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
