"""Writes the golden fixtures of the workload configurations (tests/golden/
cfg0.npz, cfg1.npz, cfg2.npz, cfg4.npz) from the FP64 OpenMP CPU oracle, in
this container.  The GPU tests (tests/test_gpu_configs.py) compare the HIP
path at the configurations' full sizes against them.

  python tests/golden/make_fixtures.py [cfg0 cfg1 cfg2 cfg4] [--threads N]

Each fixture holds list and grid digests after initialisation, the step-1
fields at a fixed sample of particles and integrals at checkpoint steps
(tests/workloads.py).  The oracle is the checker: it runs here once, and
only its outputs travel.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "dem-sphsource_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import scenes  # noqa: E402
import workloads as W  # noqa: E402

T0 = time.time()


def say(msg):
    print(f"[{time.time() - T0:8.1f}s] {msg}", flush=True)


def run_oracle(sc, steps, checkpoints, flags, dt, threads, label):
    o = oracle.Oracle(threads=threads)
    o.set_flags(flags)
    o.upload(sc)
    o.initialize(True)
    lists = W.list_digests(o)
    grids = W.grid_digests(o)
    say(f"{label}: initialised ({sc.n_fluid} fluid, {sc.n_solid} solid, dt={dt:.4e})")
    ints, fields = {}, None
    t = time.time()
    for s in range(1, steps + 1):
        o.step(s, dt, sc.gap)
        if s == 1:
            fields = W.step1_fields(o, sc)
        if s in checkpoints:
            ints[s] = W.integrals(o, sc)
            say(f"{label}: step {s} ({(time.time() - t) / s:.3f} s/step) {ints[s]}")
    o.close()
    return lists, grids, fields, ints


def make_cfg(name, threads):
    c = W.CONFIGS[name]
    sc = c["build"]()
    lists, grids, fields, ints = run_oracle(sc, c["steps"], c["checkpoints"], c["flags"], sc.dt, threads, name)
    meta = dict(n_fluid=sc.n_fluid, n_solid=sc.n_solid, dt=sc.dt, steps=c["steps"], flags=c["flags"])
    W.save(name, W.pack(lists, grids, fields, ints, c["checkpoints"], meta))
    say(f"{name}: wrote {W.fixture_path(name)}")


def make_cfg4(threads):
    scs = scenes.ensemble(256)
    dts = W.ensemble_dts(scs)
    d = dict(steps=W.CFG4_STEPS, subset=np.asarray(W.CFG4_SUBSET, np.int64))
    for k in W.CFG4_SUBSET:
        sc = scs[k]
        lists, grids, fields, ints = run_oracle(sc, W.CFG4_STEPS, W.CFG4_CHECKPOINTS, W.FLAG_ART_VISCOSITY,
                                                dts[k], threads, f"cfg4 scene {k}")
        sub = W.pack(lists, grids, fields, ints, W.CFG4_CHECKPOINTS,
                     dict(n_fluid=sc.n_fluid, n_solid=sc.n_solid, dt=dts[k]))
        for key, v in sub.items():
            d[f"s{k}_{key}"] = v
    W.save("cfg4", d)
    say(f"cfg4: wrote {W.fixture_path('cfg4')}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="*", default=["cfg0", "cfg2", "cfg4", "cfg1"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    for name in a.configs:
        if name == "cfg4":
            make_cfg4(a.threads)
        else:
            make_cfg(name, a.threads)


if __name__ == "__main__":
    main()
