"""Frames: the step loop's VTU output (dsph_frame_begin / dsph_frame_write),
the reference's solverBase::outputFluidVTU, outputSolidVTU and
outputFluid2SolidVTU (core/solverBase.cpp:69-333) as binary appended VTU.

Each written array is parsed back from the file and compared with the
library's FP64 read-back of the same state, rounded to FP32 the way the
reference's Float32 arrays hold it: bit-exact.  The snapshot is the state at
dsph_frame_begin even when the caller steps before the write, with two
slots in flight and with the write on another thread."""
import os
import re
import threading
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import dsph
import scenes

pytestmark = pytest.mark.gpu

_DT = {"Float32": np.float32, "Int32": np.int32, "UInt8": np.uint8}


def read_vtu(path):
    """A binary appended (raw, UInt64 headers) VTU file: FieldData,
    NumberOfPoints and every DataArray by name ("points" for the points)."""
    raw = open(path, "rb").read()
    tag = b'<AppendedData encoding="raw">'
    at = raw.index(tag)
    head = raw[:at].decode()
    ET.fromstring(head + "</VTKFile>")  # the header is well-formed XML
    base = raw.index(b"_", at + len(tag)) + 1
    assert raw.rstrip().endswith(b"</VTKFile>")

    def block(off, dtype, comps):
        n = int(np.frombuffer(raw, np.uint64, 1, base + off)[0])
        a = np.frombuffer(raw, dtype, n // np.dtype(dtype).itemsize, base + off + 8)
        return a.reshape(-1, comps) if comps > 1 else a

    out = {"n": int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))}
    for m in re.finditer(r"<DataArray ([^>]*)/>", head):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        comps = int(attrs.get("NumberOfComponents", 1))
        name = attrs.get("Name", "points")
        out[name] = block(int(attrs["offset"]), _DT[attrs["type"]], comps)
    return out


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
    g.frame_write(1, str(tmp_path), 2, dsph.FRAME_ALL)
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
    g.frame_begin(1, 0, 0.0)
    g.frame_write(1, str(tmp_path), 3, dsph.FRAME_SOLID)
    assert sorted(os.listdir(tmp_path)) == ["solid_0003.vtu"]
    with pytest.raises(dsph.DsphError):  # written: the slot is empty again
        g.frame_write(1, str(tmp_path), 3)
    g.frame_begin(0, 0, 0.0)
    with pytest.raises(dsph.DsphError) as e:  # a directory that does not exist
        g.frame_write(0, str(tmp_path / "missing"), 0)
    assert e.value.code == -8
    g.close()
