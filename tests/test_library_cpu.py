"""CPU-side checks of the C ABI library: it loads, it exports every entry
point include/dsph.h declares, and host-side argument checks work without a
GPU (no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import dsph
import scenes
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dsph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dsph_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_api():
    names = declared_functions()
    for must in ("dsph_create", "dsph_upload_scene", "dsph_initialize", "dsph_neighbor_search",
                 "dsph_fluid_integrate", "dsph_contact_forces", "dsph_step", "dsph_step_n",
                 "dsph_download_list", "dsph_download_grid"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = dsph.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", dsph.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (dsph_[a-z_0-9]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)
    assert set(dsph.EXPORTS) <= set(declared_functions())


def test_version_and_argument_checks_without_compute():
    lib = dsph.load_library()
    assert b"gfx950" in lib.dsph_version()
    h = C.c_void_p()
    # invalid precision is rejected before any device call
    assert lib.dsph_create(0, 7, C.byref(h)) == -1
    # null context paths
    assert lib.dsph_destroy(None) == 0
    assert lib.dsph_upload_scene(None, None) == -1


def test_scene_struct_matches_header_field_order():
    text = open(HEADER).read()
    body = text[text.index("typedef struct dsph_scene"):text.index("} dsph_scene;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([a-z_0-9]+)\s*(?:\[3\])?;", body)
    py = [f[0] for f in dsph.CScene._fields_]
    assert names == py


def test_scene_builders_match_documented_sizes():
    sc = scenes.dam_break(0.02)
    assert (sc.n_fluid, sc.n_solid) == (8000, 32256)
    big = scenes.coupled_dam_break(0.005)
    assert big.n_fluid == 512000
    assert big.n_solid == 452376 + 2000
    assert big.n_wall == 5
    assert abs(big.dt - 5.80e-5) < 1e-7
    # deterministic inputs
    again = scenes.coupled_dam_break(0.005)
    assert np.array_equal(big.solid_pos, again.solid_pos)


def test_lattice_is_the_reference_lattice():
    # getRegularPackedPoints: x outer, z inner, spacing size/num, centred
    p = scenes.regular_packed_points((0, 0, 0), (0.1, 0.2, 0.3), 0.1)
    assert p.shape == (6, 3)
    assert np.allclose(p[0], [0.05, 0.05, 0.05]) and np.allclose(p[1], [0.05, 0.05, 0.15])
    one = scenes.regular_packed_points((1, 1, 1), (0.01, 0.2, 0.3), 0.1)
    assert np.allclose(one, [[1.005, 1.1, 1.15]])


def test_cpp_host_api_builds_and_reports_errors():
    """The solverBase/dataBase C++ API and its case driver build with g++
    against the C ABI; without a GPU the driver stops with the library's
    DSPH_E_NODEVICE instead of crashing."""
    import os
    exe = os.path.join(os.path.dirname(dsph.LIB_PATH), "_build", "coupledDamBreak")
    assert os.path.exists(exe)
    r = subprocess.run([exe, "0.1", "2", "1", "1", "1"], capture_output=True, text=True, timeout=300)
    if r.returncode == 0:  # a GPU is present
        assert "coupledDamBreak:" in r.stdout
    else:
        assert r.returncode == 1 and "no HIP device" in r.stderr, r.stderr


def test_cpp_host_lattice_and_random_streams_match_python(tmp_path):
    """getRegularPackedPoints and the SplitMix64 streams of the C++ host API
    give the same bits as scenes.py."""
    import os
    src = tmp_path / "probe.cpp"
    host = os.path.join(os.path.dirname(dsph.LIB_PATH), "host")
    src.write_text(r'''
#include <cstdio>
#include "generatePoints.h"
#include "seededRandom.h"
int main() {
    auto p = getRegularPackedPoints(make_double3(-0.015, 0.1, 0.2), make_double3(0.43, 0.21, 0.37), 0.005);
    std::fwrite(p.data(), sizeof(double3), p.size(), stdout);
    auto u = uniform(kRootSeed, 1001, 100, -0.001, 0.001);
    std::fwrite(u.data(), sizeof(double), u.size(), stdout);
    return 0;
}
''')
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", host, str(src), "-o", str(exe)], check=True)
    raw = subprocess.run([str(exe)], capture_output=True, check=True).stdout
    pts = scenes.regular_packed_points((-0.015, 0.1, 0.2), (0.43, 0.21, 0.37), 0.005)
    n = pts.size
    got = np.frombuffer(raw[:8 * n], np.float64).reshape(-1, 3)
    assert np.array_equal(got, pts)
    u = scenes.uniform(scenes.ROOT_SEED, 1001, 100, -0.001, 0.001)
    assert np.array_equal(np.frombuffer(raw[8 * n:], np.float64), u)


@pytest.mark.parametrize("dx,L,W,H", [(0.02, 1.6, 0.6, 0.6), (0.0103, 1.2, 0.6, 0.6),
                                      (0.025, 1.0, 0.35, 0.4), (0.005, 0.06, 0.05, 0.05)])
def test_tank_wall_points_equal_the_filtered_lattice(dx, L, W, H):
    """The shell generator gives the points (and order) of filtering the whole
    getRegularPackedPoints lattice with the damBreak.cpp wall test."""
    o = np.array([-3 * dx] * 3)
    size = np.array([L + 6 * dx, W + 6 * dx, H + 6 * dx])
    sp0 = scenes.regular_packed_points(o, size, dx)
    keep = (sp0[:, 0] < 0) | (sp0[:, 1] < 0) | (sp0[:, 2] < 0) | (sp0[:, 0] > L) | (sp0[:, 1] > W)
    assert np.array_equal(scenes.tank_wall_points(o, size, dx, L, W), sp0[keep])
