"""ctypes loader of the FP64 OpenMP CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's CPU-baseline
leg as the checker.  The product (libdsph) never imports or calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_ROOT, "dem-sphsource_amd"))
from dsph_scene import CScene  # noqa: E402

LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    lib.orc_create.restype = C.c_void_p
    lib.orc_destroy.argtypes = [C.c_void_p]
    lib.orc_upload_scene.argtypes = [C.c_void_p, C.POINTER(CScene)]
    lib.orc_append_scene.argtypes = [C.c_void_p, C.POINTER(CScene)]
    lib.orc_append_scene.restype = C.c_int
    for name, args in [
        ("orc_set_flags", [C.c_uint32]), ("orc_set_threads", [C.c_int]),
        ("orc_initialize", [C.c_int]), ("orc_neighbor_search", [C.c_int32, C.c_int32]),
        ("orc_solid_before_contact", [C.c_double]),
        ("orc_fluid_integrate", [C.c_double, C.c_int32, C.c_int32]),
        ("orc_contact_forces", [C.c_double]), ("orc_solid_after_contact", [C.c_double]),
        ("orc_step", [C.c_int32, C.c_double, C.c_int32]),
        ("orc_add_external_force", [C.c_int32, _dp]), ("orc_add_global_damping", [C.c_int32, C.c_double]),
        ("orc_download_fluid", [_dp] * 5), ("orc_download_solid", [_dp] * 6),
        ("orc_download_clump", [_dp] * 6),
        ("orc_download_fluid_solid_force", [_dp, C.c_int64, _i64p]),
        ("orc_download_contacts", [_dp] * 5 + [C.c_int64, _i64p]),
        ("orc_download_bonds", [_dp] * 4 + [_ip, C.c_int64, _i64p]),
        ("orc_download_list", [C.c_int, _ip, _ip, C.c_int64, _i64p, _ip]),
        ("orc_download_transpose", [C.c_int, _ip, C.c_int64, _ip, _ip]),
        ("orc_grid_info", [_ip, _dp, _ip]),
        ("orc_download_grid", [C.c_int, _ip, _ip, _ip, _ip]),
    ]:
        f = getattr(lib, name)
        f.argtypes = [C.c_void_p] + args
        f.restype = C.c_int
    lib.orc_wendland.argtypes = [C.c_double, C.c_double]
    lib.orc_wendland.restype = C.c_double
    lib.orc_grad_wendland.argtypes = [_dp, C.c_double, _dp]
    lib.orc_contact.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                C.c_double, _dp, _dp, _dp]
    lib.orc_parallel_bond.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                      C.c_double, _dp]
    lib.orc_parallel_bond.restype = C.c_int
    lib.orc_rotate_inverse_inertia.argtypes = [_dp, _dp, _dp]
    lib.orc_quaternion_rotate.argtypes = [_dp, _dp, C.c_double, _dp]
    lib.orc_rotate_vector.argtypes = [_dp, _dp, _dp]
    lib.orc_combined_index.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.orc_combined_index.restype = C.c_int
    head = [C.c_char_p, C.c_double, C.c_int32, C.c_int64]
    lib.orc_vtu_fluid_ascii.argtypes = head + [_dp] * 6
    lib.orc_vtu_solid_ascii.argtypes = head + [_dp, _dp, _ip, _ip, _dp, _dp, _dp]
    lib.orc_vtu_fluid_solid_ascii.argtypes = head + [_ip, _ip, _dp, _dp, _dp]
    for f in (lib.orc_vtu_fluid_ascii, lib.orc_vtu_solid_ascii, lib.orc_vtu_fluid_solid_ascii):
        f.restype = C.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def vec(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ivec(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _vtu_ok(rc, path):
    if rc != 0:
        raise OSError(f"cannot write {path}")


# ASCII VTU frames of host arrays, as the reference's writers print them
# (core/solverBase.cpp:69-333): see vtu_ascii.cpp
def vtu_fluid_ascii(path, time, step, pos, h, vel, rho0, drho, c):
    a = [vec(pos), vec(h), vec(vel), vec(rho0), vec(drho), vec(c)]
    _vtu_ok(lib().orc_vtu_fluid_ascii(str(path).encode(), time, step, len(a[1]), *map(_p, a)), path)


def vtu_solid_ascii(path, time, step, pos, radius, cluster, clump, vel, angvel, normal):
    d = [vec(pos), vec(radius)]
    i = [_ivec(cluster), _ivec(clump)]
    v = [vec(vel), vec(angvel), vec(normal)]
    _vtu_ok(lib().orc_vtu_solid_ascii(str(path).encode(), time, step, len(d[1]), _p(d[0]), _p(d[1]),
                                      _p(i[0], _ip), _p(i[1], _ip), *map(_p, v)), path)


def vtu_fluid_solid_ascii(path, time, step, fluid, solid, fluid_pos, solid_pos, force):
    i = [_ivec(fluid), _ivec(solid)]
    d = [vec(fluid_pos), vec(solid_pos), vec(force)]
    _vtu_ok(lib().orc_vtu_fluid_solid_ascii(str(path).encode(), time, step, len(i[0]), _p(i[0], _ip),
                                            _p(i[1], _ip), *map(_p, d)), path)


class Oracle:
    """Drives the CPU oracle with the same scene and phase sequence as the
    product's C ABI."""

    def __init__(self, threads: int = 0):
        self.L = lib()
        self.h = self.L.orc_create()
        if threads:
            self.L.orc_set_threads(self.h, threads)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed: {rc}")

    def upload(self, scene):
        self.scene = scene
        cs = scene.cstruct()
        self._ok(self.L.orc_upload_scene(self.h, C.byref(cs)), "upload")
        self.nf, self.ns, self.nc = scene.n_fluid, scene.n_solid, scene.n_clump

    def append(self, extra):
        """Particles added after the run started (orc_append_scene)."""
        cs = extra.cstruct()
        self._ok(self.L.orc_append_scene(self.h, C.byref(cs)), "append")
        self.nf += extra.n_fluid
        self.ns += extra.n_solid
        self.nc += extra.n_clump

    def set_flags(self, flags):
        self.L.orc_set_flags(self.h, flags)

    def initialize(self, compute_normals=True):
        self._ok(self.L.orc_initialize(self.h, int(compute_normals)), "initialize")

    def neighbor_search(self, i_step, gap=1):
        self.L.orc_neighbor_search(self.h, i_step, gap)

    def solid_before_contact(self, dt):
        self.L.orc_solid_before_contact(self.h, dt)

    def fluid_integrate(self, dt, i_step, gap=1):
        self.L.orc_fluid_integrate(self.h, dt, i_step, gap)

    def contact_forces(self, dt):
        self.L.orc_contact_forces(self.h, dt)

    def solid_after_contact(self, dt):
        self.L.orc_solid_after_contact(self.h, dt)

    def step(self, i_step, dt, gap=1):
        self.L.orc_step(self.h, i_step, dt, gap)

    def add_external_force(self, idx, f):
        self._ok(self.L.orc_add_external_force(self.h, idx, _p(vec(f))), "add_external_force")

    def add_global_damping(self, idx, cd):
        self._ok(self.L.orc_add_global_damping(self.h, idx, float(cd)), "add_global_damping")

    def run(self, n_steps, dt, gap=1, first=1):
        for s in range(first, first + n_steps):
            self.step(s, dt, gap)

    # ---- read-back
    def fluid(self):
        n = self.nf
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        drho, sw = np.zeros(n), np.zeros(n)
        self.L.orc_download_fluid(self.h, _p(pos), _p(vel), _p(acc), _p(drho), _p(sw))
        return dict(pos=pos, vel=vel, acc=acc, drho=drho, sumw0=sw)

    def solid(self):
        n = self.ns
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "acc", "angvel", "torque", "normal")}
        self.L.orc_download_solid(self.h, *[_p(out[k]) for k in
                                            ("pos", "vel", "acc", "angvel", "torque", "normal")])
        return out

    def clump(self):
        n = self.nc
        out = {k: np.zeros((n, 3)) for k in ("centroid", "vel", "acc", "angvel", "torque")}
        out["orientation"] = np.zeros((n, 4))
        self.L.orc_download_clump(self.h, *[_p(out[k]) for k in
                                            ("centroid", "vel", "acc", "angvel", "torque",
                                             "orientation")])
        return out

    def list(self, kind):
        n = C.c_int64(0)
        self.L.orc_download_list(self.h, kind, None, None, 0, C.byref(n), None)
        m = n.value
        a, b = np.zeros(m, np.int32), np.zeros(m, np.int32)
        cnt_len = self.nf if kind in (0, 1) else self.ns
        cnt = np.zeros(max(cnt_len, 1), np.int32)
        self.L.orc_download_list(self.h, kind, _p(a, _ip), _p(b, _ip), m, C.byref(n),
                                 _p(cnt, _ip) if kind != 3 else None)
        return a, b, cnt[:cnt_len]

    def transpose(self, kind):
        a, _, _ = self.list(kind)
        m = len(a)
        npt = self.nf if kind == 0 else self.ns
        idx = np.zeros(max(m, 1), np.int32)
        rs, re = np.zeros(max(npt, 1), np.int32), np.zeros(max(npt, 1), np.int32)
        self.L.orc_download_transpose(self.h, kind, _p(idx, _ip), m, _p(rs, _ip), _p(re, _ip))
        return idx[:m], rs[:npt], re[:npt]

    def fs_force(self):
        n = C.c_int64(0)
        self.L.orc_download_fluid_solid_force(self.h, None, 0, C.byref(n))
        f = np.zeros((n.value, 3))
        self.L.orc_download_fluid_solid_force(self.h, _p(f), n.value, C.byref(n))
        return f

    def contacts(self):
        n = C.c_int64(0)
        self.L.orc_download_contacts(self.h, None, None, None, None, None, 0, C.byref(n))
        m = n.value
        out = {k: np.zeros((m, 3)) for k in ("force", "torque", "sliding", "rolling", "torsion")}
        self.L.orc_download_contacts(self.h, *[_p(out[k]) for k in
                                               ("force", "torque", "sliding", "rolling", "torsion")],
                                     m, C.byref(n))
        return out

    def bonds(self):
        n = C.c_int64(0)
        self.L.orc_download_bonds(self.h, None, None, None, None, None, 0, C.byref(n))
        m = n.value
        fn, tt = np.zeros(m), np.zeros(m)
        fs, tb = np.zeros((m, 3)), np.zeros((m, 3))
        on = np.zeros(m, np.int32)
        self.L.orc_download_bonds(self.h, _p(fn), _p(tt), _p(fs), _p(tb), _p(on, _ip), m,
                                  C.byref(n))
        return dict(normal_force=fn, torsion=tt, shear=fs, bending=tb, is_bonded=on)

    def grid_info(self):
        gs = np.zeros(3, np.int32)
        cs = np.zeros(3)
        nc = C.c_int32(0)
        self.L.orc_grid_info(self.h, _p(gs, _ip), _p(cs), C.byref(nc))
        return gs, cs, nc.value

    def grid(self, species):
        gs, cs, ncell = self.grid_info()
        n = self.nf if species == 0 else self.ns
        h, p = np.zeros(n, np.int32), np.zeros(n, np.int32)
        s, e = np.zeros(ncell, np.int32), np.zeros(ncell, np.int32)
        self.L.orc_download_grid(self.h, species, _p(h, _ip), _p(p, _ip), _p(s, _ip), _p(e, _ip))
        return dict(hash=h, perm=p, start=s, end=e)


# ---- single formulas ------------------------------------------------------
def wendland(r, h):
    return lib().orc_wendland(float(r), float(h))


def grad_wendland(rij, h):
    out = np.zeros(3)
    lib().orc_grad_wendland(_p(vec(rij)), float(h), _p(out))
    return out


def contact(model, v_rel, w_rel, n, overlap, m_eff, r_eff, dt, params, springs=None):
    sp = vec(np.zeros(9) if springs is None else np.asarray(springs, float).reshape(9))
    out = np.zeros(6)
    lib().orc_contact(model, _p(vec(v_rel)), _p(vec(w_rel)), _p(vec(n)), overlap, m_eff, r_eff,
                      dt, _p(vec(params)), _p(sp), _p(out))
    return out[:3], out[3:], sp.reshape(3, 3)


def parallel_bond(state, n_prev, n, v_rel, w_a, w_b, r_a, r_b, dt, params):
    st = vec(np.asarray(state, float).copy())
    on = lib().orc_parallel_bond(_p(st), _p(vec(n_prev)), _p(vec(n)), _p(vec(v_rel)),
                                 _p(vec(w_a)), _p(vec(w_b)), r_a, r_b, dt, _p(vec(params)))
    return st, on


def rotate_inverse_inertia(q, inv_i):
    out = np.zeros(6)
    lib().orc_rotate_inverse_inertia(_p(vec(q)), _p(vec(inv_i)), _p(out))
    return out


def quaternion_rotate(q, w, dt):
    out = np.zeros(4)
    lib().orc_quaternion_rotate(_p(vec(q)), _p(vec(w)), dt, _p(out))
    return out


def rotate_vector(v, ang):
    out = np.zeros(3)
    lib().orc_rotate_vector(_p(vec(v)), _p(vec(ang)), _p(out))
    return out


def combined_index(n, a, b):
    return lib().orc_combined_index(n, a, b)
