"""The workload configurations at their stated sizes on the HIP path, against
golden fixtures the FP64 CPU oracle wrote (tests/golden/make_fixtures.py,
tests/workloads.py):

- lists and hashed grids after initialisation bit-exact (sha256 digests);
- the fields after step 1 within 1e-10 (FP64 path) and 1e-5 (mixed path),
  norm-wise relative, at a fixed sample of particles;
- whole-scene integrals within 1% at every checkpoint.

cfg3 (16M fluid + 20,000 spheres) has no oracle fixture: it runs the FP64
and the mixed GPU paths side by side and asserts the mixed path's per-field
error against FP64 after 1, 100 and 1000 steps; the table is written to
gpurun_out/precision_study_16m.md.
"""
import json
import os
import sys
import time

import numpy as np
import pytest

import dsph
import refmodel
import scenes
import workloads as W
from conftest import ROOT

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-5
TOL_INTEGRAL = 0.01
T0 = time.time()


def note(msg):
    print(f"[{time.time() - T0:7.1f}s] {msg}", flush=True)
    sys.stdout.flush()


def check_initial(g, fx, kinds=(0, 1, 2, 3)):
    got = W.list_digests(g, kinds)
    for kind in kinds:
        assert got[kind][0] == int(fx[f"list{kind}_n"]), kind
        assert got[kind][1] == str(fx[f"list{kind}_sha"]), kind
    grids = W.grid_digests(g)
    for sp in (0, 1):
        assert grids[sp] == str(fx[f"grid{sp}_sha"]), sp


def check_step1(g, sc, fx, tol):
    got = W.step1_fields(g, sc)
    worst = {}
    for k, v in got.items():
        ref = fx["step1_" + k]
        assert v.shape == ref.shape, k
        worst[k] = refmodel.rel_err(v, ref, floor=1e-300)
    bad = {k: e for k, e in worst.items() if e > tol}
    assert not bad, (bad, worst)
    return worst


def run_checkpoints(g, sc, fx, dt, checkpoints, gap=1, ranges=None):
    """Graph-replayed steps to every checkpoint; integrals against the
    fixture's at 1%."""
    ref = W.unpack_integrals(fx)
    done = 0
    worst = 0.0
    for cp in checkpoints:
        g.run(cp - done, dt, gap, first=done + 1, graphs=True)
        done = cp
        got = W.integrals(g, sc, *(ranges or ()))
        dev = W.integral_deviations(got, ref[cp])
        worst = max(worst, max(dev.values()))
        assert max(dev.values()) < TOL_INTEGRAL, (cp, {k: v for k, v in dev.items() if v >= TOL_INTEGRAL},
                                                  got, ref[cp])
    return worst


def solver(sc, prec, flags=0):
    g = dsph.Solver(prec)
    g.set_flags(flags)
    g.upload(sc)
    g.initialize()
    return g


# --------------------------------------------------------- cfg0, cfg1, cfg2
_BUILT = {}


def built(name):
    if name not in _BUILT:
        _BUILT.clear()
        _BUILT[name] = (W.CONFIGS[name]["build"](), W.load(name))
    return _BUILT[name]


@pytest.mark.parametrize("name", ["cfg0", "cfg2", "cfg1"])
@pytest.mark.parametrize("prec", [dsph.PREC_FP64, dsph.PREC_MIXED])
def test_config_at_stated_size_against_oracle_fixture(name, prec):
    sc, fx = built(name)
    c = W.CONFIGS[name]
    assert (sc.n_fluid, sc.n_solid) == (int(fx["n_fluid"]), int(fx["n_solid"]))
    note(f"{name} prec={prec}: {sc.n_fluid} fluid + {sc.n_solid} solid, {c['steps']} steps")
    g = solver(sc, prec, c["flags"])
    check_initial(g, fx)
    note(f"{name}: lists and grids bit-exact")
    g.step(1, sc.dt, sc.gap)
    w = check_step1(g, sc, fx, TOL_FP64 if prec == dsph.PREC_FP64 else TOL_MIXED)
    note(f"{name}: step-1 fields worst {max(w.values()):.2e}")
    g.close()
    # the checkpoint run from a fresh context, every step from HIP graphs
    g = solver(sc, prec, c["flags"])
    worst = run_checkpoints(g, sc, fx, sc.dt, c["checkpoints"], sc.gap)
    note(f"{name}: integrals at {c['checkpoints']} worst deviation {worst:.2e}")
    g.close()


# -------------------------------------------------------------------- cfg4
@pytest.fixture(scope="module")
def ensemble():
    t = time.time()
    scs = scenes.ensemble(256)
    note(f"cfg4: 256 scenes built in {time.time() - t:.1f}s, "
         f"{sum(s.n_fluid for s in scs)} fluid + {sum(s.n_solid for s in scs)} solid")
    return scs, W.load("cfg4")


def test_ensemble_scenes_alone_match_the_oracle(ensemble):
    scs, fx = ensemble
    for k in W.CFG4_SUBSET:
        sc = scs[k]
        sub = {key[len(f"s{k}_"):]: v for key, v in fx.items() if key.startswith(f"s{k}_")}
        g = solver(sc, dsph.PREC_MIXED, W.FLAG_ART_VISCOSITY)
        check_initial(g, sub)
        assert float(sub["dt"]) == sc.dt
        g.step(1, sc.dt, sc.gap)
        check_step1(g, sc, sub, TOL_MIXED)
        g.close()
        note(f"cfg4 scene {k} alone: lists bit-exact, step 1 within {TOL_MIXED}")


def test_ensemble_batch_of_256_against_the_oracle(ensemble):
    scs, fx = ensemble
    g = dsph.Solver(dsph.PREC_MIXED)
    g.set_flags(W.FLAG_ART_VISCOSITY)
    g.upload_batch(scs)
    # every scene at its own time step, as the oracle runs it alone
    g.set_scene_time_steps(W.ensemble_dts(scs))
    g.initialize()
    info = g.batch_info()
    note("cfg4: batch uploaded and initialised")
    refs = {k: W.unpack_integrals({key[len(f"s{k}_"):]: v for key, v in fx.items()
                                   if key.startswith(f"s{k}_")}) for k in W.CFG4_SUBSET}
    done, worst = 0, 0.0
    for cp in W.CFG4_CHECKPOINTS:
        g.run(cp - done, 0.0, first=done + 1, graphs=True)
        done = cp
        for k in W.CFG4_SUBSET:
            got = W.integrals(g, scs[k], (int(info["fluid"][k]), int(info["fluid"][k + 1])),
                              (int(info["solid"][k]), int(info["solid"][k + 1])))
            dev = W.integral_deviations(got, refs[k][cp])
            worst = max(worst, max(dev.values()))
            assert max(dev.values()) < TOL_INTEGRAL, (k, cp, dev)
    note(f"cfg4: 256-scene batch, {W.CFG4_STEPS} steps, subset integrals worst {worst:.2e}")
    g.close()


# -------------------------------------------------------------------- cfg3
def field_errors(a, b, sc, sa, sb):
    """Mixed (a) against FP64 (b): norm-wise relative errors of density
    change, velocity, pressure and acceleration, and maxima over the
    physical scales (rho0, sqrt(2 g H), rho0 g H); DEM position over the
    sphere radius."""
    rho0 = float(np.asarray(sc.fluid_rho0)[0])
    c = float(np.asarray(sc.fluid_c)[0])
    H = 0.4
    vs = np.sqrt(2 * 9.81 * H)
    e = dict(
        rho_rel_l2=refmodel.rel_err(rho0 + a["drho"], rho0 + b["drho"]),
        drho_max_over_rho0=float(np.max(np.abs(a["drho"] - b["drho"]))) / rho0,
        p_rel_l2=refmodel.rel_err(a["drho"], b["drho"]),
        v_rel_l2=refmodel.rel_err(a["vel"], b["vel"]),
        v_max_over_vs=float(np.max(np.linalg.norm(a["vel"] - b["vel"], axis=1))) / vs,
        p_max_over_rho0gH=c * c * float(np.max(np.abs(a["drho"] - b["drho"]))) / (rho0 * 9.81 * H),
        acc_rel_l2=refmodel.rel_err(a["acc"], b["acc"]),
        x_max_over_h=float(np.max(np.linalg.norm(a["pos"] - b["pos"], axis=1))) / float(sc.fluid_h[0]),
    )
    mov = W.moving(sc)
    r = float(np.asarray(sc.solid_radius)[mov[0]])
    e["dem_x_max_over_r"] = float(np.max(np.linalg.norm(sa["pos"][mov] - sb["pos"][mov], axis=1))) / r
    e["dem_v_rel_l2"] = refmodel.rel_err(sa["vel"][mov], sb["vel"][mov])
    return e


# bars of the mixed path against FP64: the one-step bar on every field
# after step 1, the integral bar (1%) at every checkpoint
CFG3_BARS = {
    1: dict(rho_rel_l2=1e-5, p_rel_l2=1e-5, v_rel_l2=1e-5, acc_rel_l2=1e-5, dem_x_max_over_r=1e-5),
}


def test_precision_study_16m_fp64_against_mixed():
    t = time.time()
    sc = W.cfg3_scene()
    note(f"cfg3: {sc.n_fluid} fluid + {sc.n_solid} solid built in {time.time() - t:.1f}s, dt={sc.dt:.4e}")
    assert sc.n_fluid == 15_625_000 and int((np.asarray(sc.solid_inv_mass) > 0).sum()) == 20_000
    g64 = solver(sc, dsph.PREC_FP64)
    note("cfg3: FP64 context ready")
    gmx = solver(sc, dsph.PREC_MIXED)
    note("cfg3: mixed context ready")
    table, ints, done = {}, {}, 0
    for cp in (1, 100, 1000):
        t = time.time()
        for g in (g64, gmx):
            g.run(cp - done, sc.dt, first=done + 1, graphs=True)
            g.synchronize()
        done = cp
        fa, fb = gmx.fluid(), g64.fluid()
        sa, sb = gmx.solid(), g64.solid()
        table[cp] = field_errors(fa, fb, sc, sa, sb)
        del fa, fb, sa, sb
        ia, ib = W.integrals(gmx, sc), W.integrals(g64, sc)
        ints[cp] = W.integral_deviations(ia, ib)
        note(f"cfg3 step {cp} ({time.time() - t:.1f}s): " +
             ", ".join(f"{k}={v:.2e}" for k, v in table[cp].items()))
    rb = {"fp64": int(g64.status()[dsph.ST_REBUILDS]), "mixed": int(gmx.status()[dsph.ST_REBUILDS])}
    g64.close()
    gmx.close()
    out = os.path.join(ROOT, "gpurun_out")
    os.makedirs(out, exist_ok=True)
    keys = list(table[1])
    with open(os.path.join(out, "precision_study_16m.md"), "w") as fh:
        fh.write(f"# Precision study: {sc.n_fluid} fluid + {sc.n_solid} solid (dx = 0.0016), "
                 f"dt = {sc.dt:.4e}\n\nMixed path (FP32 pair terms summed in FP64, FP64 "
                 f"integration) against the all-FP64 GPU path, same scene, same steps "
                 f"(tests/test_gpu_configs.py::test_precision_study_16m_fp64_against_mixed).\n"
                 f"Superset rebuilds over the run: {rb}.\n\n")
        fh.write("| step | " + " | ".join(keys) + " | worst integral |\n")
        fh.write("|---" * (len(keys) + 2) + "|\n")
        for cp in table:
            fh.write(f"| {cp} | " + " | ".join(f"{table[cp][k]:.2e}" for k in keys) +
                     f" | {max(ints[cp].values()):.2e} |\n")
        fh.write("\nBars asserted: " + json.dumps(CFG3_BARS) + "; integrals within 1%.\n")
    for cp, bars in CFG3_BARS.items():
        for k, bar in bars.items():
            assert table[cp][k] < bar, (cp, k, table[cp][k], bar)
    for cp in table:
        assert max(ints[cp].values()) < TOL_INTEGRAL, (cp, ints[cp])
