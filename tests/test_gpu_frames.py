"""Frames: the step loop's VTU output (dsph_frame_begin / dsph_frame_write),
the reference's solverBase::outputFluidVTU, outputSolidVTU and
outputFluid2SolidVTU (core/solverBase.cpp:69-333) as binary appended VTU.

Each written array is parsed back from the file and compared with the
library's FP64 read-back of the same state, rounded to FP32 the way the
reference's Float32 arrays hold it: bit-exact.  The snapshot is the state at
dsph_frame_begin even when the caller steps before the write, with two
slots in flight and with the write on another thread.  Against the
reference's own ASCII files of the same state (the oracle's restatement of
its writers), every value agrees to the precision those files print."""
import os
import threading

import numpy as np
import pytest

import dsph
import scenes
import oracle
from vtu import read_vtu, read_vtu_ascii

pytestmark = pytest.mark.gpu


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32)


def expected(g, sc):
    """The frame's arrays from the FP64 read-back (caller order)."""
    fl, so = g.fluid(), g.solid()
    fi, sj, _ = g.list(dsph.LIST_FLUID_SOLID)
    F = g.fs_force()
    c = np.asarray(sc.fluid_c, np.float64)
    rho0 = np.asarray(sc.fluid_rho0, np.float64)
    fluid = {"points": f32(fl["pos"]), "smooth length": f32(sc.fluid_h), "velocity": f32(fl["vel"]),
             "density": f32(rho0 + fl["drho"]), "pressure": f32(fl["drho"] * (c * c))}
    ns = sc.n_solid
    solid = {"points": f32(so["pos"]), "radius": f32(sc.solid_radius),
             "clusterID": np.asarray(sc.solid_cluster if sc.solid_cluster is not None else -np.ones(ns), np.int32),
             "clumpID": np.asarray(sc.solid_clump if sc.solid_clump is not None else -np.ones(ns), np.int32),
             "velocity": f32(so["vel"]), "angularVelocity": f32(so["angvel"]), "normal": f32(so["normal"])}
    pairs = {(int(a), int(b)): (f32(0.5 * (fl["pos"][a] + so["pos"][b])), f32(F[k]))
             for k, (a, b) in enumerate(zip(fi, sj))}
    return fluid, solid, pairs


def check_frame(d, frame, step, time, exp):
    fluid, solid, pairs = exp
    num = f"{frame:04d}"
    for name, want in (("fluid", fluid), ("solid", solid)):
        got = read_vtu(os.path.join(d, f"{name}_{num}.vtu"))
        assert got["n"] == len(want["points"])
        assert got["TIME"][0] == np.float32(time) and got["STEP"][0] == step
        for k, v in want.items():
            assert np.array_equal(got[k], v), (name, k)
    got = read_vtu(os.path.join(d, f"fluid2Solid_{num}.vtu"))
    assert got["n"] == len(pairs) and len(pairs) > 0
    keys = list(zip(got["fluidID"].tolist(), got["solidID"].tolist()))
    assert set(keys) == set(pairs)  # the same pairs (in the device's order)
    for k, key in enumerate(keys):
        mid, force = pairs[key]
        assert np.array_equal(got["points"][k], mid) and np.array_equal(got["force"][k], force), key


SCENE = scenes.immersed_spheres()


@pytest.mark.parametrize("prec", [dsph.PREC_FP64, dsph.PREC_MIXED])
def test_frame_arrays_match_the_state_at_begin(prec, tmp_path):
    sc = SCENE
    g = dsph.Solver(prec)
    g.upload(sc)
    g.initialize()
    g.run(3, sc.dt, sc.gap)
    g.frame_begin(0, 3, 3 * sc.dt)
    exp = expected(g, sc)
    g.run(2, sc.dt, sc.gap, first=4)  # the copy holds step 3, not these
    g.frame_write(0, str(tmp_path), 7)
    check_frame(str(tmp_path), 7, 3, 3 * sc.dt, exp)
    g.close()


def test_two_slots_and_a_writer_thread(tmp_path):
    sc = SCENE
    g = dsph.Solver(dsph.PREC_MIXED)
    g.upload(sc)
    g.initialize()
    g.run(2, sc.dt, sc.gap, graphs=True)
    g.frame_begin(0, 2, 2 * sc.dt)
    exp0 = expected(g, sc)
    g.run(3, sc.dt, sc.gap, first=3, graphs=True)
    g.frame_begin(1, 5, 5 * sc.dt)
    exp1 = expected(g, sc)
    # slot 0 is written on another thread while this one keeps stepping
    t = threading.Thread(target=g.frame_write, args=(0, str(tmp_path), 1))
    t.start()
    g.run(4, sc.dt, sc.gap, first=6, graphs=True)
    t.join()
    g.frame_write(1, str(tmp_path), 2)
    check_frame(str(tmp_path), 1, 2, 2 * sc.dt, exp0)
    check_frame(str(tmp_path), 2, 5, 5 * sc.dt, exp1)
    g.close()


def test_frame_species_and_misuse(tmp_path):
    sc = SCENE
    g = dsph.Solver(dsph.PREC_MIXED)
    g.upload(sc)
    g.initialize()
    with pytest.raises(dsph.DsphError) as e:  # nothing begun on the slot
        g.frame_write(0, str(tmp_path), 0)
    assert e.value.code == -4
    with pytest.raises(dsph.DsphError):
        g.frame_begin(2, 0, 0.0)  # slots 0 and 1 only
    for bad in (0, 8):  # no species, an unknown one
        with pytest.raises(dsph.DsphError):
            g.frame_begin(1, 0, 0.0, bad)
    g.frame_begin(1, 0, 0.0, dsph.FRAME_SOLID)
    g.frame_write(1, str(tmp_path), 3)
    assert sorted(os.listdir(tmp_path)) == ["solid_0003.vtu"]
    with pytest.raises(dsph.DsphError):  # written: the slot is empty again
        g.frame_write(1, str(tmp_path), 3)
    g.frame_begin(0, 0, 0.0)
    with pytest.raises(dsph.DsphError) as e:  # a directory that does not exist
        g.frame_write(0, str(tmp_path / "missing"), 0)
    assert e.value.code == -8
    g.close()


def _ascii_bound(b):
    """|binary - ascii| for one value x: the binary file holds fl32(x) (half
    an FP32 ulp from x), the ASCII file x to 10 decimals (5e-11 from x)."""
    b = np.abs(np.asarray(b, np.float32))
    return 0.5 * np.spacing(b).astype(np.float64) + 5e-11 * (1 + 1e-6)


@pytest.mark.parametrize("prec", [dsph.PREC_FP64, dsph.PREC_MIXED])
def test_frame_agrees_with_the_reference_ascii_files(prec, tmp_path):
    """The same state written both ways: the binary frame and the reference's
    ASCII files (printed from the FP64 read-back, the reference's host
    mirrors).  Points, arrays and TIME agree within the two roundings, the
    integer arrays and STEP exactly, pairs matched by their fluid and solid
    indices."""
    sc = SCENE
    g = dsph.Solver(prec)
    g.upload(sc)
    g.initialize()
    g.run(4, sc.dt, sc.gap)
    time = 4 * sc.dt
    g.frame_begin(0, 4, time)
    g.frame_write(0, str(tmp_path), 0)
    fl, so = g.fluid(), g.solid()
    fi, sj, _ = g.list(dsph.LIST_FLUID_SOLID)
    F = g.fs_force()
    ns = sc.n_solid
    clus = sc.solid_cluster if sc.solid_cluster is not None else -np.ones(ns)
    clump = sc.solid_clump if sc.solid_clump is not None else -np.ones(ns)
    ref = tmp_path / "ref"
    ref.mkdir()
    oracle.vtu_fluid_ascii(ref / "fluid_0000.vtu", time, 4, fl["pos"], sc.fluid_h, fl["vel"],
                           sc.fluid_rho0, fl["drho"], sc.fluid_c)
    oracle.vtu_solid_ascii(ref / "solid_0000.vtu", time, 4, so["pos"], sc.solid_radius, clus, clump,
                           so["vel"], so["angvel"], so["normal"])
    oracle.vtu_fluid_solid_ascii(ref / "fluid2Solid_0000.vtu", time, 4, fi, sj, fl["pos"], so["pos"], F)
    g.close()
    for name, arrays in (("fluid", ["points", "smooth length", "velocity", "density", "pressure"]),
                         ("solid", ["points", "radius", "clusterID", "clumpID", "velocity",
                                    "angularVelocity", "normal"])):
        b = read_vtu(str(tmp_path / f"{name}_0000.vtu"))
        a = read_vtu_ascii(str(ref / f"{name}_0000.vtu"))
        assert b["n"] == a["n"] > 0 and b["STEP"][0] == a["STEP"][0] == 4
        assert abs(b["TIME"][0] - a["TIME"][0]) <= _ascii_bound(b["TIME"])[0]
        for k in ["connectivity", "offsets", "types"]:
            assert np.array_equal(b[k], a[k]), (name, k)
        for k in arrays:
            if b[k].dtype == np.int32:
                assert np.array_equal(b[k], a[k]), (name, k)
            else:
                err = np.abs(b[k].astype(np.float64) - a[k])
                assert np.all(err <= _ascii_bound(b[k])), (name, k, float(err.max()))
    b = read_vtu(str(tmp_path / "fluid2Solid_0000.vtu"))
    a = read_vtu_ascii(str(ref / "fluid2Solid_0000.vtu"))
    assert b["n"] == a["n"] == len(fi) > 0
    at = {(int(x), int(y)): k for k, (x, y) in enumerate(zip(b["fluidID"], b["solidID"]))}
    order = np.array([at[(int(x), int(y))] for x, y in zip(fi, sj)])
    for k in ("points", "force"):
        err = np.abs(b[k][order].astype(np.float64) - a[k])
        assert np.all(err <= _ascii_bound(b[k][order])), (k, float(err.max()))
