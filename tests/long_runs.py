"""Long validation runs of the HIP path at full workload sizes (not collected
by pytest: each takes minutes).

  python tests/long_runs.py integrals --workload coupled --steps 1000 --out r.json
      the 1M-particle coupled dam break (dx = 0.005: 512,000 fluid,
      452,376 wall dummies, 2,000 Hertz-Mindlin spheres) run for N steps on
      the GPU (mixed and FP64) and on the FP64 OpenMP oracle; per-field
      errors after step 1 and the integrals (fluid mass, kinetic and
      potential energy, surge front, mean density, DEM centroid) at each
      checkpoint, against the 1% bar of the multi-step comparison.
  python tests/long_runs.py integrals --workload dam_break --steps 2000
      the same for the 40k-particle plain dam break (dx = 0.02).
  python tests/long_runs.py precision --dx 0.0016
      the 16M-fluid dam break with 20,000 DEM spheres (50x20x20 lattice at
      0.022 m from (0.5, 0.09, 0.012), jitter U(+-0.05 r)): GPU FP64 against
      GPU mixed, per-field errors at steps 1, 100 and 1000; optionally one
      oracle step against both.

The oracle is the checker here, never the thing measured.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(_ROOT, "dem-sphsource_amd"), os.path.join(_ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import dsph  # noqa: E402
import oracle  # noqa: E402
import scenes  # noqa: E402

TOL_INTEGRAL = 0.01
T0 = time.time()


def say(msg):
    print(f"[{time.time() - T0:7.1f}s] {msg}", flush=True)


def integrals(f, s, sc):
    """Whole-scene integrals from downloaded fields: fluid mass from the
    particle volumes (h = 1.3 dx), kinetic and potential energy, surge front,
    mean density, and the centroid of the moving DEM particles."""
    m = (np.asarray(sc.fluid_h) / 1.3) ** 3 * np.asarray(sc.fluid_rho0)
    rho = np.asarray(sc.fluid_rho0) + f["drho"]
    out = dict(mass=float(np.sum(m)),
               volume=float(np.sum(m / rho)),
               ke=float(0.5 * np.sum(m * np.sum(f["vel"] ** 2, axis=1))),
               pe=float(np.sum(m * 9.81 * f["pos"][:, 2])),
               front=float(np.max(f["pos"][:, 0])),
               mean_rho=float(np.mean(rho)))
    if sc.n_solid and s is not None:
        mov = np.asarray(sc.solid_inv_mass) > 0
        if mov.any():
            out["dem_centroid_x"] = float(np.mean(s["pos"][mov, 0]))
            out["dem_centroid_z"] = float(np.mean(s["pos"][mov, 2]))
            ms = 1.0 / np.asarray(sc.solid_inv_mass)[mov]
            out["dem_ke"] = float(0.5 * np.sum(ms * np.sum(s["vel"][mov] ** 2, axis=1)))
    return out


def rel_dev(a, b, floor):
    return abs(a - b) / max(abs(b), floor)


def integral_floors(ref):
    """Scale below which a relative error is not meaningful: kinetic energies
    start at zero, so they are compared against 1e-6 of the potential energy
    (about the kinetic energy after 1 ms of free fall of the column)."""
    fl = {k: 1e-12 for k in ref}
    fl["ke"] = 1e-6 * abs(ref["pe"])
    if "dem_ke" in ref:
        fl["dem_ke"] = 1e-6 * abs(ref["pe"])
    return fl


def field_errors(a, b, sc, sa=None, sb=None):
    """Per-field deviations of state a from state b (caller order):
    density as max |d rho| / rho0, velocity and position as max norm over the
    fluid's velocity scale sqrt(2 g H) and smoothing length, pressure
    (c^2 d rho) over the hydrostatic scale rho0 g H, DEM position over the
    sphere radius."""
    rho0 = np.asarray(sc.fluid_rho0)
    c = np.asarray(sc.fluid_c)
    H = 0.4
    vs = np.sqrt(2 * 9.81 * H)
    out = dict(
        rho_max=float(np.max(np.abs(a["drho"] - b["drho"]) / rho0)),
        v_max=float(np.max(np.linalg.norm(a["vel"] - b["vel"], axis=1)) / vs),
        v_rel_l2=float(np.linalg.norm(a["vel"] - b["vel"]) / max(np.linalg.norm(b["vel"]), 1e-300)),
        x_max_over_h=float(np.max(np.linalg.norm(a["pos"] - b["pos"], axis=1)
                                  / np.asarray(sc.fluid_h))),
        p_max=float(np.max(np.abs(c * c * (a["drho"] - b["drho"]))) / (rho0.max() * 9.81 * H)),
        acc_rel_l2=float(np.linalg.norm(a["acc"] - b["acc"]) / max(np.linalg.norm(b["acc"]), 1e-300)),
    )
    if sa is not None and sb is not None and sc.n_solid:
        mov = np.asarray(sc.solid_inv_mass) > 0
        if mov.any():
            r = np.asarray(sc.solid_radius)[mov]
            d = np.linalg.norm(sa["pos"][mov] - sb["pos"][mov], axis=1) / r
            out["dem_x_max_over_r"] = float(np.max(d))
            dv = np.linalg.norm(sa["vel"][mov] - sb["vel"][mov], axis=1)
            out["dem_v_max"] = float(np.max(dv) / vs)
    return out


def gpu_run(sc, precision, checkpoints, keep_fields=()):
    """Runs the scene on the GPU in graph-captured steps up to each
    checkpoint; returns integrals per checkpoint and full fields at the steps
    listed in keep_fields."""
    t = time.time()
    g = dsph.Solver(precision)
    g.upload(sc)
    g.initialize(True)
    say(f"gpu prec={precision} initialised in {time.time() - t:.1f}s")
    res, fields, done = {}, {}, 0
    for cp in checkpoints:
        t = time.time()
        g.run(cp - done, sc.dt, first=done + 1, graphs=True)
        g.synchronize()
        dt_run = time.time() - t
        f = g.fluid()
        s = g.solid() if sc.n_solid else None
        res[cp] = integrals(f, s, sc)
        if cp in keep_fields:
            fields[cp] = (f, s)
        say(f"gpu prec={precision} step {cp}: {dt_run:.2f}s for {cp - done} steps; "
            f"ke={res[cp]['ke']:.6e} front={res[cp]['front']:.5f}")
        done = cp
    g.close()
    return res, fields


def oracle_run(sc, checkpoints, keep_fields=(), progress_every=25):
    t = time.time()
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize(True)
    say(f"oracle initialised in {time.time() - t:.1f}s")
    res, fields, done = {}, {}, 0
    t_all = time.time()
    for cp in checkpoints:
        for s in range(done + 1, cp + 1):
            o.step(s, sc.dt, sc.gap)
            if s % progress_every == 0:
                el = time.time() - t_all
                say(f"oracle step {s}/{checkpoints[-1]} ({el / s:.2f}s/step)")
        f = o.fluid()
        so = o.solid() if sc.n_solid else None
        res[cp] = integrals(f, so, sc)
        if cp in keep_fields:
            fields[cp] = (f, so)
        say(f"oracle step {cp}: ke={res[cp]['ke']:.6e} front={res[cp]['front']:.5f}")
        done = cp
    o.close()
    return res, fields


def cmd_integrals(args):
    if args.workload == "coupled":
        sc = scenes.coupled_dam_break(dx=args.dx or 0.005)
    else:
        sc = scenes.dam_break(dx=args.dx or 0.02)
    n = sc.n_fluid + sc.n_solid
    say(f"{sc.name}: {sc.n_fluid} fluid + {sc.n_solid} solid ({n} particles), dt={sc.dt:.4e}, "
        f"{args.steps} steps")
    cps = sorted({1, *[c for c in (10, 100, 250, 500, 1000, 1500, 2000) if c < args.steps],
                  args.steps})
    out = dict(workload=sc.name, n_fluid=sc.n_fluid, n_solid=sc.n_solid, dt=sc.dt,
               steps=args.steps, checkpoints=cps, tol=TOL_INTEGRAL, runs={})
    runs = {}
    for name, prec in (("mixed", dsph.PREC_MIXED), ("fp64", dsph.PREC_FP64)):
        runs[name] = gpu_run(sc, prec, cps, keep_fields=(1, args.steps))
    if not args.no_oracle:
        runs["oracle"] = oracle_run(sc, cps, keep_fields=(1, args.steps))
    ref_name = "oracle" if "oracle" in runs else "fp64"
    ref, ref_f = runs[ref_name]
    worst = 0.0
    for name, (res, flds) in runs.items():
        rec = dict(integrals={str(k): v for k, v in res.items()})
        if name != ref_name:
            dev = {}
            for cp in cps:
                fl = integral_floors(ref[cp])
                dev[str(cp)] = {k: rel_dev(res[cp][k], ref[cp][k], fl[k]) for k in ref[cp]}
            rec["rel_dev_vs_" + ref_name] = dev
            w = max(dev[str(cps[-1])].values())
            rec["worst_final"] = w
            worst = max(worst, w)
            for cp in (1, args.steps):
                fa, sa = flds[cp]
                fb, sb = ref_f[cp]
                rec[f"fields_step{cp}_vs_{ref_name}"] = field_errors(fa, fb, sc, sa, sb)
        out["runs"][name] = rec
        say(f"{name}: " + ("reference" if name == ref_name else
                           f"worst final integral deviation {rec['worst_final']:.3e}"))
    out["pass"] = bool(worst < TOL_INTEGRAL)
    say(f"integrals within {TOL_INTEGRAL}: {out['pass']} (worst {worst:.3e})")
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(out, fh, indent=1)
    return 0 if out["pass"] else 1


def cfg3_scene(dx):
    return scenes.coupled_dam_break(dx=dx, n_dem=(50, 20, 20), dem_spacing=0.022,
                                    dem_origin=(0.5, 0.09, 0.012), jitter=0.05, config_id=3,
                                    name="coupled_dam_break_16m")


def cmd_precision(args):
    t = time.time()
    sc = cfg3_scene(args.dx)
    say(f"{sc.name}: {sc.n_fluid} fluid + {sc.n_solid} solid, dt={sc.dt:.4e} "
        f"(scene built in {time.time() - t:.1f}s)")
    cps = [c for c in (1, 100, 1000) if c <= args.steps]
    runs = {}
    for name, prec in (("mixed", dsph.PREC_MIXED), ("fp64", dsph.PREC_FP64)):
        runs[name] = gpu_run(sc, prec, cps, keep_fields=tuple(cps))
    out = dict(workload=sc.name, n_fluid=sc.n_fluid, n_solid=sc.n_solid, dt=sc.dt, checkpoints=cps,
               mixed_vs_fp64={}, integrals={})
    for cp in cps:
        fa, sa = runs["mixed"][1][cp]
        fb, sb = runs["fp64"][1][cp]
        e = field_errors(fa, fb, sc, sa, sb)
        out["mixed_vs_fp64"][str(cp)] = e
        fl = integral_floors(runs["fp64"][0][cp])
        out["integrals"][str(cp)] = {k: rel_dev(runs["mixed"][0][cp][k], v, fl[k])
                                     for k, v in runs["fp64"][0][cp].items()}
        say(f"step {cp}: mixed vs fp64 " + ", ".join(f"{k}={v:.2e}" for k, v in e.items()))
    if args.oracle_steps:
        ores, oflds = oracle_run(sc, [1], keep_fields=(1,), progress_every=1)
        fo, so = oflds[1]
        for name in ("mixed", "fp64"):
            fa, sa = runs[name][1][1]
            e = field_errors(fa, fo, sc, sa, so)
            out[f"{name}_vs_oracle_step1"] = e
            say(f"step 1: {name} vs oracle " + ", ".join(f"{k}={v:.2e}" for k, v in e.items()))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(out, fh, indent=1)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("integrals")
    a.add_argument("--workload", choices=["coupled", "dam_break"], default="coupled")
    a.add_argument("--dx", type=float, default=None)
    a.add_argument("--steps", type=int, default=1000)
    a.add_argument("--no-oracle", action="store_true")
    a.add_argument("--out", default=None)
    b = sub.add_parser("precision")
    b.add_argument("--dx", type=float, default=0.0016)
    b.add_argument("--steps", type=int, default=1000)
    b.add_argument("--oracle-steps", type=int, default=0, help="1: also one oracle step")
    b.add_argument("--out", default=None)
    args = ap.parse_args()
    return cmd_integrals(args) if args.cmd == "integrals" else cmd_precision(args)


if __name__ == "__main__":
    sys.exit(main())
