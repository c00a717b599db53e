"""Full-state checkpoints (dsph_save_state / dsph_load_state) and the
roll-back of a graph-replayed run whose lists overflow (dsph_step_n).

A run saved after N steps and continued in a fresh context from the blob
must give bit for bit the state of the same run continued without the
interruption: fluid, solids, clumps, contact springs, bonds, lists."""
import numpy as np
import pytest

import dsph
import scenes

pytestmark = pytest.mark.gpu


def snapshot(g):
    out = {}
    for name, d in (("f", g.fluid()), ("s", g.solid()), ("c", g.clump())):
        for k, v in d.items():
            out[f"{name}_{k}"] = v
    for k, v in g.contacts().items():
        out["contact_" + k] = v
    for k, v in g.bonds().items():
        out["bond_" + k] = v
    out["fs_force"] = g.fs_force()
    for kind in (dsph.LIST_FLUID_FLUID, dsph.LIST_FLUID_SOLID, dsph.LIST_SOLID_SOLID, dsph.LIST_BOND):
        a, b, c = g.list(kind)
        out[f"list{kind}"] = np.concatenate([a, b, c])
    return out


def same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


CASES = {
    "coupled": lambda: scenes.coupled_dam_break(dx=0.02, n_dem=(4, 3, 3), dem_spacing=0.025,
                                                dem_origin=(0.45, 0.2, 0.012), name="coupled_small"),
    "granular": lambda: scenes.granular_column(n_sphere=600, n_clump=60, n_bonded=216),
    "dam_break_case": scenes.dam_break_case_small,
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", [dsph.PREC_FP64, dsph.PREC_MIXED])
def test_checkpoint_resume_is_bit_identical(case, prec):
    sc = CASES[case]()
    n1, n2 = 2 * sc.gap + 7, 2 * sc.gap + 9
    a = dsph.Solver(prec)
    a.upload(sc)
    a.set_skin(0.2)
    a.set_tuning(dsph.TUNE_GROUP, 2)
    a.initialize()
    a.run(n1, sc.dt, sc.gap, graphs=True)
    blob = a.save_state()
    a.run(n2, sc.dt, sc.gap, first=n1 + 1, graphs=True)
    ref = snapshot(a)
    a.close()
    b = dsph.Solver(prec)
    b.upload(sc)                  # the same scene, not initialised
    b.load_state(blob)            # takes the run's skin, tuning and capacities
    assert b.get_tuning(dsph.TUNE_GROUP) == 2
    b.run(n2, sc.dt, sc.gap, first=n1 + 1, graphs=True)
    same(snapshot(b), ref)
    # and eagerly, in the context that saved it
    c = dsph.Solver(prec)
    c.upload(sc)
    c.initialize()
    c.load_state(blob)
    for s in range(n1 + 1, n1 + n2 + 1):
        c.step(s, sc.dt, sc.gap)
    same(snapshot(c), ref)


def test_load_state_rejects_another_scene():
    sc = CASES["coupled"]()
    a = dsph.Solver(dsph.PREC_FP64)
    a.upload(sc)
    a.initialize()
    a.run(3, sc.dt)
    blob = a.save_state()
    other = scenes.dam_break(0.04)
    b = dsph.Solver(dsph.PREC_FP64)
    b.upload(other)
    with pytest.raises(dsph.DsphError) as e:
        b.load_state(blob)
    assert e.value.code == -1
    m = dsph.Solver(dsph.PREC_MIXED)
    m.upload(sc)
    with pytest.raises(dsph.DsphError):
        m.load_state(blob)
    with pytest.raises(dsph.DsphError):
        m.load_state(blob[:100])


@pytest.mark.parametrize("prec", [dsph.PREC_FP64, dsph.PREC_MIXED])
def test_step_n_rolls_back_an_overflowing_run(prec):
    """Capacities far too small for the run: every list overflows inside a
    graph-replayed run.  The run is rolled back and stepped again eagerly
    with growing lists, and ends where eager checked steps end."""
    sc = CASES["coupled"]()
    ref = dsph.Solver(prec)
    ref.upload(sc)
    ref.set_capacity(4, 16, 4)
    ref.set_superset_capacity(8, 2)
    ref.initialize()
    for s in range(1, 41):
        ref.step(s, sc.dt)
    g = dsph.Solver(prec)
    g.upload(sc)
    g.initialize()
    g.run(20, sc.dt, graphs=True)
    g.set_capacity(4, 16, 4)          # shrink mid-run: the next replay overflows
    g.set_superset_capacity(8, 2)
    g.run(20, sc.dt, first=21, graphs=True)
    st = g.status()
    assert st[dsph.ST_FF_OVERFLOW] == st[dsph.ST_FS_OVERFLOW] == st[dsph.ST_SS_OVERFLOW] == 0
    g.run(5, sc.dt, first=41, graphs=True)   # later replays work
    ref.run(5, sc.dt, first=41, graphs=True)
    a, b = snapshot(g), snapshot(ref)
    for k in ("f_pos", "f_vel", "s_pos", "s_vel", "list0", "list1", "list2"):
        if prec == dsph.PREC_FP64:
            assert np.allclose(a[k], b[k], rtol=1e-9, atol=1e-12), k
        else:
            assert np.allclose(a[k], b[k], rtol=1e-5, atol=1e-9), k
