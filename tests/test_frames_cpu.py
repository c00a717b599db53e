"""The oracle's ASCII VTU writers (oracle/vtu_ascii.cpp), the reference's
frame format of core/solverBase.cpp:69-333, pinned on small hand-made
arrays: the element order the reference prints, every Float32 value as fixed
text with 10 decimals (solverBase.cpp:81), the derived arrays (density =
rho0 + dRho, pressure = dRho c^2, pair points at the midpoint) against
values worked out by hand, and the empty frame."""
import re
import xml.etree.ElementTree as ET

import numpy as np

import oracle
from vtu import read_vtu_ascii

FIXED10 = re.compile(r"-?\d+\.\d{10}$")


def _close(a, b):
    # printed to 10 decimals: within half a unit of the 10th decimal
    return np.all(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) <= 5e-11 * (1 + 1e-9))


def test_fluid_frame_values_and_layout(tmp_path):
    p = tmp_path / "fluid_0003.vtu"
    pos = np.array([[0.0, 1.25, -2.5], [1e3, 1e-12, 0.1234567890123]])
    vel = np.array([[1.0, -1.0, 0.5], [0.0, 2.0, -3.0]])
    h = np.array([0.013, 0.026])
    rho0 = np.array([1000.0, 1.2])
    drho = np.array([2.5, -0.1])
    c = np.array([20.0, 3.0])
    oracle.vtu_fluid_ascii(p, 0.75, 12, pos, h, vel, rho0, drho, c)
    a = read_vtu_ascii(str(p))
    assert a["order"] == ["TIME", "STEP", "points", "connectivity", "offsets", "types", "smooth length",
                          "velocity", "density", "pressure"]
    assert a["n"] == 2 and a["STEP"][0] == 12 and _close(a["TIME"], [0.75])
    assert all(FIXED10.match(t) for t in a["tokens"])  # TIME, then the points
    assert a["tokens"][1:4] == ["0.0000000000", "1.2500000000", "-2.5000000000"]
    assert _close(a["points"], pos) and _close(a["velocity"], vel) and _close(a["smooth length"], h)
    assert _close(a["density"], [1002.5, 1.1])
    assert _close(a["pressure"], [2.5 * 400.0, -0.1 * 9.0])
    assert a["connectivity"].tolist() == [0, 1] and a["offsets"].tolist() == [1, 2]
    assert a["types"].tolist() == [1, 1]
    root = ET.parse(str(p)).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("version") == "0.1"
    assert root.find("UnstructuredGrid/Piece/PointData").get("Scalars") == "smooth length"


def test_solid_frame_values_and_integer_arrays(tmp_path):
    p = tmp_path / "solid_0000.vtu"
    n = 3
    pos = np.arange(9, dtype=np.float64).reshape(3, 3) * 0.1
    rad = np.array([0.01, 0.02, 0.03])
    clus = np.array([-1, 0, 7])
    clump = np.array([-1, -1, 2])
    vel = -pos
    w = pos + 1
    nrm = np.array([[0, 0, 1.0], [1.0, 0, 0], [0, 0, 0]])
    oracle.vtu_solid_ascii(p, 0.0, 0, pos, rad, clus, clump, vel, w, nrm)
    a = read_vtu_ascii(str(p))
    assert a["order"] == ["TIME", "STEP", "points", "connectivity", "offsets", "types", "radius", "clusterID",
                          "clumpID", "velocity", "angularVelocity", "normal"]
    assert a["n"] == n
    assert a["clusterID"].tolist() == [-1, 0, 7] and a["clumpID"].tolist() == [-1, -1, 2]
    for k, v in (("points", pos), ("radius", rad), ("velocity", vel), ("angularVelocity", w), ("normal", nrm)):
        assert _close(a[k], v), k
    root = ET.parse(str(p)).getroot()
    assert root.find("UnstructuredGrid/Piece/PointData").get("Scalars") == "radius"


def test_pair_frame_points_at_the_midpoints(tmp_path):
    p = tmp_path / "fluid2Solid_0001.vtu"
    fpos = np.array([[0.0, 0.0, 0.0], [2.0, 2.0, 2.0]])
    spos = np.array([[1.0, 2.0, 3.0], [-1.0, 0.0, 1.0]])
    fluid = np.array([0, 1, 1])
    solid = np.array([0, 0, 1])
    force = np.array([[1.0, 0, 0], [0, -2.0, 0], [0, 0, 3.5]])
    oracle.vtu_fluid_solid_ascii(p, 1.0, 5, fluid, solid, fpos, spos, force)
    a = read_vtu_ascii(str(p))
    assert a["order"] == ["TIME", "STEP", "points", "connectivity", "offsets", "types", "force"]
    assert _close(a["points"], [[0.5, 1.0, 1.5], [1.5, 2.0, 2.5], [0.5, 1.0, 1.5]])
    assert _close(a["force"], force)


def test_empty_frames_parse(tmp_path):
    e = np.zeros((0, 3))
    z = np.zeros(0)
    oracle.vtu_fluid_ascii(tmp_path / "f.vtu", 0.0, 0, e, z, e, z, z, z)
    oracle.vtu_fluid_solid_ascii(tmp_path / "p.vtu", 0.0, 0, np.zeros(0, np.int32), np.zeros(0, np.int32),
                                 e, e, e)
    for f in ("f.vtu", "p.vtu"):
        a = read_vtu_ascii(str(tmp_path / f))
        assert a["n"] == 0 and a["points"].size == 0
