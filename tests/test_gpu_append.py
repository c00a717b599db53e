"""Particles added after the run started (dsph_append_scene), the
reference's dataBase::addFluid / addCluster / addClump called mid-run
(core/dataBase.cpp:113-117, 226-346), HIP path against the oracle's
orc_append_scene: the state right after the addition (lists, bonds, kept
springs, the new particles' kernel sums) and after more steps, FP64 and
mixed; an empty addition continues the run; misuse is refused."""
import sys

import numpy as np
import pytest

import dsph
import oracle
import refmodel
import scenes

pytestmark = pytest.mark.gpu

TOL = {dsph.PREC_FP64: 1e-10, dsph.PREC_MIXED: 1e-5}
BEFORE, AFTER = 6, 30


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b, floor=1e-300):
    return refmodel.rel_err(np.asarray(a), np.asarray(b), floor=floor)


@pytest.mark.parametrize("prec", [dsph.PREC_FP64, dsph.PREC_MIXED])
def test_addition_matches_the_oracle(prec):
    note(f"poured scene, addition after {BEFORE} steps, prec={prec}")
    base, extra = scenes.poured()
    g = dsph.Solver(prec)
    g.upload(base)
    g.initialize()
    o = oracle.Oracle()
    o.upload(base)
    o.initialize()
    g.run(BEFORE, base.dt)
    o.run(BEFORE, base.dt)
    kept = (g.fluid(), g.solid(), g.clump())
    g.append(extra)
    o.append(extra)
    # every particle that was there keeps its state, bit for bit
    for was, now, n in zip(kept, (g.fluid(), g.solid(), g.clump()),
                           (base.n_fluid, base.n_solid, base.n_clump)):
        for k in was:
            assert np.array_equal(now[k][:n], was[k]), k
    # against the oracle: FP64 at its one-step bar; the mixed path carries
    # the FP32 error of the steps before the addition (the bar of its
    # multi-step comparisons)
    tol = TOL[prec] if prec == dsph.PREC_FP64 else 20 * TOL[prec]
    # right after: the lists of the step-0 search, the bonds, the springs
    # carried over, the new fluid's kernel sums
    for kind in (dsph.LIST_FLUID_FLUID, dsph.LIST_FLUID_SOLID, dsph.LIST_SOLID_SOLID, dsph.LIST_BOND):
        a, b = g.list(kind), o.list(kind)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), kind
    gb, ob = g.bonds(), o.bonds()
    assert len(gb["is_bonded"]) == 16
    for k in ob:  # bond forces and moments: rounding noise below 1 nN
        assert rel(gb[k], ob[k], 1e-9) < tol, k
    gc, oc = g.contacts(), o.contacts()
    for k in ("sliding", "rolling", "torsion"):
        assert rel(gc[k], oc[k], 1e-12) < tol, k
    assert np.abs(oc["sliding"]).max() > 0
    gf, of = g.fluid(), o.fluid()
    assert rel(gf["sumw0"], of["sumw0"]) < tol
    for k in ("pos", "vel", "drho"):
        assert rel(gf[k], of[k], 1e-3 if k != "pos" else 0.0) < tol, k
    # then the run goes on from there (replayed from HIP graphs)
    g.run(AFTER, base.dt, first=BEFORE + 1, graphs=True)
    o.run(AFTER, base.dt, first=BEFORE + 1)
    gf, of = g.fluid(), o.fluid()
    for k in ("pos", "vel", "drho"):
        assert rel(gf[k], of[k], 1e-3 if k != "pos" else 0.0) < 20 * TOL[prec], k
    gs, os_ = g.solid(), o.solid()
    for k in ("pos", "vel", "angvel"):
        assert rel(gs[k], os_[k], 1e-6 if k != "pos" else 0.0) < 20 * TOL[prec], k
    gk, ok_ = g.clump(), o.clump()
    for k in ("centroid", "vel"):
        assert rel(gk[k], ok_[k], 1e-6 if k != "centroid" else 0.0) < 20 * TOL[prec], k
    g.close()


def test_an_empty_addition_continues_the_run():
    base, _ = scenes.poured()
    empty = scenes.empty_scene(base.domain_origin, base.domain_size, n_material=base.n_material)
    runs = []
    for add in (False, True):
        g = dsph.Solver(dsph.PREC_FP64)
        g.upload(base)
        g.initialize()
        g.run(BEFORE, base.dt)
        if add:
            g.append(empty)
        g.run(AFTER, base.dt, first=BEFORE + 1)
        runs.append((g.fluid(), g.solid(), g.clump(), g.bonds()))
        g.close()
    # the fluid's slots are sorted again at the addition, so sums run in
    # another order: FP64 rounding apart
    for a, b in zip(runs[0], runs[1]):
        for k in a:
            assert a[k].shape == b[k].shape, k
            if a[k].size:
                assert rel(a[k], b[k], 1e-6) < 1e-12, k


def test_additions_are_checked():
    base, extra = scenes.poured()
    g = dsph.Solver(dsph.PREC_MIXED)
    g.upload(base)
    with pytest.raises(dsph.DsphError) as e:  # before the initialisation
        g.append(extra)
    assert e.value.code == dsph.E_STATE
    g.initialize()
    walled = scenes.empty_scene(base.domain_origin, base.domain_size, n_material=base.n_material)
    scenes.add_infinite_wall(walled, (0, 0, 0), (0, 0, 1))
    with pytest.raises(dsph.DsphError) as e:
        g.append(walled)
    assert e.value.code == dsph.E_INVALID
    g.close()
    b = dsph.Solver(dsph.PREC_MIXED)
    b.upload_batch([base, base])
    b.initialize()
    with pytest.raises(dsph.DsphError) as e:  # a batch
        b.append(extra)
    assert e.value.code == dsph.E_STATE
    b.close()
