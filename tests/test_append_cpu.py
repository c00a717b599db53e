"""Particles added after the run started: the oracle's orc_append_scene,
its restatement of dataBase::addFluid / addCluster / addClump called
mid-run (core/dataBase.cpp:113-117, 226-268) and of addBondData's later
branch (:309-343).  Pinned by what the operation must leave alone (an
empty addition, or one far from everything, changes no particle that was
there, bit for bit), by the kernel sum a new particle starts from, and by
the bond rule on a hand-built pair of clusters."""
import numpy as np

import dsph
import oracle
import scenes

STEPS = 6


def fresh(sc):
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize()
    return o


def state(o, nf=None, ns=None, nc=None):
    f, s, c = o.fluid(), o.solid(), o.clump()
    out = {}
    for name, d, n in (("fluid", f, nf), ("solid", s, ns), ("clump", c, nc)):
        for k, v in d.items():
            out[f"{name}.{k}"] = v[:n] if n is not None else v
    return out


def far_block(sc):
    """Fluid in the middle of the tank, high above the fluid and the cluster
    (further than 2h from every particle, walls included)."""
    e = scenes.empty_scene(sc.domain_origin, sc.domain_size, n_material=sc.n_material)
    dx = 0.005
    pts = scenes.regular_packed_points((0.015, 0.01, 0.066), (2 * dx, 2 * dx, dx), dx)
    scenes.add_fluid(e, pts, 1.3 * dx, scenes.RHO0, scenes.sound_speed(6 * dx), scenes.NU)
    return e


def test_an_empty_addition_changes_nothing():
    base, _ = scenes.poured()
    empty = scenes.empty_scene(base.domain_origin, base.domain_size, n_material=base.n_material)
    a, b = fresh(base), fresh(base)
    a.run(STEPS, base.dt)
    b.run(STEPS, base.dt)
    a.append(empty)
    a.run(STEPS, base.dt, first=STEPS + 1)
    b.run(STEPS, base.dt, first=STEPS + 1)
    sa, sb = state(a), state(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for k in ("sliding", "rolling", "torsion"):
        assert np.array_equal(a.contacts()[k], b.contacts()[k]), k
    for k, v in a.bonds().items():
        assert np.array_equal(v, b.bonds()[k]), k


def test_a_far_addition_leaves_the_scene_alone_and_starts_from_its_kernel_sum():
    base, _ = scenes.poured()
    far = far_block(base)
    a, b = fresh(base), fresh(base)
    a.run(STEPS, base.dt)
    b.run(STEPS, base.dt)
    a.append(far)
    # a new particle's sumW0 is its kernel sum at the append: here that of
    # the block alone, which a fresh run computes at its step 1
    alone = scenes.empty_scene(base.domain_origin, base.domain_size, n_material=base.n_material)
    scenes.add_fluid(alone, far.fluid_pos, far.fluid_h[0], far.fluid_rho0[0], far.fluid_c[0], far.fluid_nu[0])
    o = fresh(alone)
    o.run(1, base.dt)
    assert np.array_equal(a.fluid()["sumw0"][base.n_fluid:], o.fluid()["sumw0"])
    assert np.all(o.fluid()["sumw0"] > 0)
    a.run(STEPS, base.dt, first=STEPS + 1)
    b.run(STEPS, base.dt, first=STEPS + 1)
    sa, sb = state(a, base.n_fluid, base.n_solid, base.n_clump), state(b)
    for k in sb:
        assert np.array_equal(sa[k], sb[k]), k
    for k, v in b.bonds().items():
        assert np.array_equal(a.bonds()[k], v), k


def test_added_cluster_bonds_and_the_springs_kept():
    base, extra = scenes.poured()
    o = fresh(base)
    o.run(STEPS, base.dt)
    before = o.bonds()
    bi0, bj0, _ = o.list(dsph.LIST_BOND)
    si0, sj0, _ = o.list(dsph.LIST_SOLID_SOLID)
    springs0 = o.contacts()["sliding"]
    assert len(bi0) == 12  # the 2 x 2 x 2 cluster: its 12 touching edges
    o.append(extra)
    after = o.bonds()
    bi, bj, _ = o.list(dsph.LIST_BOND)
    ns0 = base.n_solid
    # the old bonds first, as they were; then the 4 edges of the new 2 x 2
    # cluster, whose pointed solids all come after the last bonded one
    assert np.array_equal(bi[:12], bi0) and np.array_equal(bj[:12], bj0)
    for k, v in before.items():
        assert np.array_equal(after[k][:12], v), k
    assert len(bi) == 16 and np.all(bi[12:] >= ns0) and np.all(bj[12:] >= ns0)
    assert np.all(after["is_bonded"][12:] == 1) and np.all(after["normal_force"][12:] == 0)
    # the two clusters touch (different clusters: contacts, no bond), and the
    # contacts that were there keep their springs
    si, sj, _ = o.list(dsph.LIST_SOLID_SOLID)
    pairs = {(int(i), int(j)) for i, j in zip(si, sj)}
    assert any((i < ns0) != (j < ns0) for i, j in pairs)
    springs = o.contacts()["sliding"]
    at = {(int(i), int(j)): k for k, (i, j) in enumerate(zip(si, sj))}
    kept = [k for k, (i, j) in enumerate(zip(si0, sj0)) if (int(i), int(j)) in at]
    assert len(kept) > 0 and np.any(springs0[kept] != 0)
    for k in kept:
        assert np.array_equal(springs[at[(int(si0[k]), int(sj0[k]))]], springs0[k])


def test_additions_are_checked():
    base, extra = scenes.poured()
    o = oracle.Oracle()
    o.upload(base)
    try:
        o.append(extra)  # before the initialisation
        raise AssertionError("accepted")
    except RuntimeError:
        pass
    o.initialize()
    walled = scenes.empty_scene(base.domain_origin, base.domain_size, n_material=base.n_material)
    scenes.add_infinite_wall(walled, (0, 0, 0), (0, 0, 1))
    try:
        o.append(walled)
        raise AssertionError("accepted")
    except RuntimeError:
        pass
