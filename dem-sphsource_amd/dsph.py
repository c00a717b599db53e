"""ctypes binding of libdsph.so — the MI355X DEM–SPH hot path behind include/dsph.h.

``Solver`` mirrors the reference's ``solverBase`` driving loop
(core/solverBase.cpp:3-67): upload a scene (dataBase::buildDeviceData),
``initialize`` (first neighbour search, bonds, solid normals), then ``step``
per timestep (neighborSearch -> solidIntegrateBeforeContact -> fluidIntegrate
-> calSolidContactAfterFluidIntegrate -> solidIntegrateAfterContact).

There is no fallback: if the HIP library is not built or no GPU is present,
constructing a ``Solver`` raises.
"""
from __future__ import annotations

import ctypes as C
import time
import os

import numpy as np

from dsph_scene import CScene, Scene  # noqa: F401  (re-exported)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsph.so")

PREC_FP64, PREC_MIXED = 0, 1
FLAG_WALL_CONTINUE, FLAG_CLUMP_QUATERNION, FLAG_ART_VISCOSITY = 1, 2, 4
TUNE_PAIR_BLOCK, TUNE_BUILD_BLOCK, TUNE_GROUP, TUNE_TILE = 1, 2, 3, 4
LIST_FLUID_FLUID, LIST_FLUID_SOLID, LIST_SOLID_SOLID, LIST_BOND = 0, 1, 2, 3
ST_FF_OVERFLOW, ST_FS_OVERFLOW, ST_SS_OVERFLOW, ST_OUT_OF_DOMAIN = 0, 1, 2, 3
ST_FF_MAX, ST_FS_TOTAL, ST_SS_TOTAL, ST_SUP_OVERFLOW, ST_REBUILDS = 4, 5, 6, 7, 8
ST_SUP_MAX, ST_NONFINITE = 9, 10
STATUS_WORDS = 12
E_INVALID, E_HIP, E_CAPACITY, E_STATE, E_NOMEM, E_NODEVICE, E_NONFINITE = -1, -2, -3, -4, -5, -6, -7

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> argument types after the context pointer (all return int)
_SIGS = {
    "dsph_upload_scene": [C.POINTER(CScene)],
    "dsph_upload_batch": [C.POINTER(CScene), C.c_int32],
    "dsph_batch_info": [_ip] * 6,
    "dsph_set_capacity": [C.c_int32, C.c_int64, C.c_int64],
    "dsph_set_flags": [C.c_uint32],
    "dsph_set_gravity": [_dp],
    "dsph_set_reorder_period": [C.c_int32],
    "dsph_set_skin": [C.c_double],
    "dsph_set_tuning": [C.c_int32, C.c_int32],
    "dsph_get_tuning": [C.c_int32, _ip],
    "dsph_set_superset_capacity": [C.c_int32, C.c_int32],
    "dsph_initialize": [C.c_int],
    "dsph_neighbor_search": [C.c_int32, C.c_int32],
    "dsph_solid_before_contact": [C.c_double],
    "dsph_fluid_integrate": [C.c_double, C.c_int32, C.c_int32],
    "dsph_contact_forces": [C.c_double],
    "dsph_solid_after_contact": [C.c_double],
    "dsph_step": [C.c_int32, C.c_double, C.c_int32, C.c_int],
    "dsph_step_n": [C.c_int32, C.c_int32, C.c_double, C.c_int32],
    "dsph_synchronize": [],
    "dsph_save_state": [_vp, C.c_int64, _i64p],
    "dsph_load_state": [_vp, C.c_int64],
    "dsph_set_contact_models": [_dp, _dp, _dp],
    "dsph_set_clump_velocity": [C.c_int32, _dp],
    "dsph_add_external_force": [C.c_int32, _dp],
    "dsph_add_global_damping": [C.c_int32, C.c_double],
    "dsph_download_fluid": [_dp] * 5,
    "dsph_download_solid": [_dp] * 6,
    "dsph_download_clump": [_dp] * 6,
    "dsph_download_fluid_solid_force": [_dp, C.c_int64, _i64p],
    "dsph_append_scene": [C.POINTER(CScene)],
    "dsph_frame_begin": [C.c_int32, C.c_int32, C.c_double, C.c_uint32],
    "dsph_frame_write": [C.c_int32, C.c_char_p, C.c_int32],
    "dsph_download_contacts": [_dp] * 5 + [C.c_int64, _i64p],
    "dsph_download_bonds": [_dp] * 4 + [_ip, C.c_int64, _i64p],
    "dsph_download_list": [C.c_int, _ip, _ip, C.c_int64, _i64p, _ip],
    "dsph_grid_info": [_ip, _dp, _ip],
    "dsph_download_grid": [C.c_int, _ip, _ip, _ip, _ip],
    "dsph_status": [_i64p],
    "dsph_superset_stats": [_i64p],
    "dsph_tile_stats": [_i64p],
    "dsph_set_scene_time_steps": [_dp],
    "dsph_hash_sort": [_vp, C.c_int32, _dp, _dp, _ip, _vp, _vp, _vp, _vp],
    "dsph_hash_sort_noalloc": [_vp, C.c_int32, _dp, _dp, _ip, _vp, _vp, _vp, _vp],
    "dsph_set_timing": [C.c_int],
    "dsph_kernel_times": [C.c_char_p, C.c_int64, _dp, C.c_int32],
}

EXPORTS = ["dsph_create", "dsph_destroy", "dsph_last_error", "dsph_version"] + list(_SIGS)

FRAME_FLUID, FRAME_SOLID, FRAME_FLUID_SOLID, FRAME_ALL = 1, 2, 4, 7
FRAME_SLOTS = 2


class DsphError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with code {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libdsph.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libdsph.so is not built ({path}); run `make -C dem-sphsource_amd`")
    lib = C.CDLL(path)
    lib.dsph_create.argtypes = [C.c_int, C.c_int, C.POINTER(_vp)]
    lib.dsph_create.restype = C.c_int
    lib.dsph_destroy.argtypes = [_vp]
    lib.dsph_destroy.restype = C.c_int
    lib.dsph_last_error.argtypes = [_vp]
    lib.dsph_last_error.restype = C.c_char_p
    lib.dsph_version.restype = C.c_char_p
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = [_vp] + args
        f.restype = C.c_int
    _lib = lib
    return lib


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


class Solver:
    """One device context running one scene."""

    def __init__(self, precision: int = PREC_MIXED, device: int = 0):
        self.L = load_library()
        h = _vp()
        rc = self.L.dsph_create(device, precision, C.byref(h))
        if rc != 0:
            raise DsphError("dsph_create", rc, "no usable HIP device" if rc == -6 else "")
        self.h = h
        self.precision = precision
        self.nf = self.ns = self.nc = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.dsph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        rc = getattr(self.L, name)(self.h, *args)
        if rc < 0:
            raise DsphError(name, rc, (self.L.dsph_last_error(self.h) or b"").decode())
        return rc

    # ---- set-up
    def upload(self, scene: Scene):
        self.scene = scene
        cs = scene.cstruct()
        self._call("dsph_upload_scene", C.byref(cs))
        self.nf, self.ns, self.nc = scene.n_fluid, scene.n_solid, scene.n_clump

    def append(self, extra: Scene):
        """Particles added after the run started (dsph_append_scene): extra's
        fluid, solids and clumps go after the existing ones, which keep
        their state."""
        cs = extra.cstruct()
        self._call("dsph_append_scene", C.byref(cs))
        self.nf += extra.n_fluid
        self.ns += extra.n_solid
        self.nc += extra.n_clump

    def upload_batch(self, scenes):
        """Independent scenes advanced together, each on its own grid; their
        particles are numbered in scene order (dsph_upload_batch)."""
        self.scenes = list(scenes)
        # the structs own the numpy buffers their pointers refer to: keep them
        # alive until the library has copied the scenes
        structs = [sc.cstruct() for sc in self.scenes]
        arr = (CScene * len(structs))(*structs)
        self._call("dsph_upload_batch", arr, len(structs))
        del structs
        self.scene = self.scenes[0]
        self.nf = sum(sc.n_fluid for sc in self.scenes)
        self.ns = sum(sc.n_solid for sc in self.scenes)
        self.nc = sum(sc.n_clump for sc in self.scenes)

    def batch_info(self):
        n = C.c_int32(0)
        self._call("dsph_batch_info", C.byref(n), None, None, None, None, None)
        k = n.value
        out = {key: np.zeros(k + 1, np.int32) for key in ("fluid", "solid", "clump", "wall", "cell")}
        self._call("dsph_batch_info", C.byref(n), _p(out["fluid"], _ip), _p(out["solid"], _ip),
                   _p(out["clump"], _ip), _p(out["wall"], _ip), _p(out["cell"], _ip))
        return out

    def set_flags(self, flags):
        self._call("dsph_set_flags", flags)

    def set_tuning(self, key, value):
        self._call("dsph_set_tuning", key, int(value))

    def get_tuning(self, key):
        v = C.c_int32(0)
        self._call("dsph_get_tuning", key, C.byref(v))
        return v.value

    def apply_tuning(self, cfg):
        """Apply a configuration returned by autotune()."""
        if "pair_block" in cfg:
            self.set_tuning(TUNE_PAIR_BLOCK, cfg["pair_block"])
        if "build_block" in cfg:
            self.set_tuning(TUNE_BUILD_BLOCK, cfg["build_block"])
        if "group" in cfg:
            self.set_tuning(TUNE_GROUP, cfg["group"])
        if "tile" in cfg:
            self.set_tuning(TUNE_TILE, cfg["tile"])
        if "skin" in cfg:
            self.set_skin(cfg["skin"])

    def set_capacity(self, ff=0, fs=0, ss=0):
        self._call("dsph_set_capacity", ff, fs, ss)

    def set_reorder_period(self, p):
        self._call("dsph_set_reorder_period", p)

    def set_skin(self, skin_over_h):
        """Verlet skin of the fluid superset, in units of the largest h."""
        self._call("dsph_set_skin", float(skin_over_h))

    def set_superset_capacity(self, ff=0, fs=0):
        self._call("dsph_set_superset_capacity", ff, fs)

    def initialize(self, compute_normals=True):
        self._call("dsph_initialize", int(compute_normals))

    # ---- phases
    def neighbor_search(self, i_step, gap=1):
        self._call("dsph_neighbor_search", i_step, gap)

    def solid_before_contact(self, dt):
        self._call("dsph_solid_before_contact", dt)

    def fluid_integrate(self, dt, i_step, gap=1):
        self._call("dsph_fluid_integrate", dt, i_step, gap)

    def contact_forces(self, dt):
        self._call("dsph_contact_forces", dt)

    def solid_after_contact(self, dt):
        self._call("dsph_solid_after_contact", dt)

    def step(self, i_step, dt, gap=1, check_capacity=True):
        self._call("dsph_step", i_step, dt, gap, int(check_capacity))

    def step_n(self, i_step0, n, dt, gap=1):
        self._call("dsph_step_n", i_step0, n, dt, gap)

    def run(self, n_steps, dt, gap=1, first=1, graphs=False):
        if graphs:
            self.step_n(first, n_steps, dt, gap)
        else:
            for s in range(first, first + n_steps):
                self.step(s, dt, gap)

    def synchronize(self):
        self._call("dsph_synchronize")

    def save_state(self) -> bytes:
        """The full state as a checkpoint blob (dsph_save_state)."""
        n = C.c_int64(0)
        self._call("dsph_save_state", None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value)
        self._call("dsph_save_state", buf, n.value, C.byref(n))
        return buf.raw[: n.value]

    def load_state(self, blob: bytes):
        """Continue a saved run (dsph_load_state): the scene must be uploaded."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._call("dsph_load_state", buf, len(blob))

    def set_contact_models(self, hertz=None, linear=None, bonded=None):
        """Replace contact tables during the run (field-major, as in Scene)."""
        def arr(a):
            return None if a is None else np.ascontiguousarray(np.asarray(a, float).reshape(-1))
        h, l, b = arr(hertz), arr(linear), arr(bonded)
        self._call("dsph_set_contact_models", _p(h) if h is not None else None,
                   _p(l) if l is not None else None, _p(b) if b is not None else None)

    def set_clump_velocity(self, idx, v):
        self._call("dsph_set_clump_velocity", idx, _p(np.asarray(v, float)))

    def add_external_force(self, idx, f):
        self._call("dsph_add_external_force", idx, _p(np.asarray(f, float)))

    def add_global_damping(self, idx, cd):
        self._call("dsph_add_global_damping", idx, cd)

    # ---- read-back (caller's original order)
    def fluid(self):
        n = self.nf
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        drho, sw = np.zeros(n), np.zeros(n)
        self._call("dsph_download_fluid", _p(pos), _p(vel), _p(acc), _p(drho), _p(sw))
        return dict(pos=pos, vel=vel, acc=acc, drho=drho, sumw0=sw)

    def solid(self):
        n = self.ns
        keys = ("pos", "vel", "acc", "angvel", "torque", "normal")
        out = {k: np.zeros((n, 3)) for k in keys}
        self._call("dsph_download_solid", *[_p(out[k]) for k in keys])
        return out

    def clump(self):
        n = self.nc
        keys = ("centroid", "vel", "acc", "angvel", "torque")
        out = {k: np.zeros((n, 3)) for k in keys}
        out["orientation"] = np.zeros((n, 4))
        self._call("dsph_download_clump", *[_p(out[k]) for k in keys], _p(out["orientation"]))
        return out

    def list(self, kind):
        n = C.c_int64(0)
        self._call("dsph_download_list", kind, None, None, 0, C.byref(n), None)
        m = n.value
        a, b = np.zeros(max(m, 1), np.int32), np.zeros(max(m, 1), np.int32)
        cnt_len = self.nf if kind in (0, 1) else self.ns
        cnt = np.zeros(max(cnt_len, 1), np.int32)
        self._call("dsph_download_list", kind, _p(a, _ip), _p(b, _ip), m, C.byref(n),
                   _p(cnt, _ip) if kind != LIST_BOND else None)
        return a[:m], b[:m], cnt[:cnt_len]

    def fs_force(self):
        n = C.c_int64(0)
        self._call("dsph_download_fluid_solid_force", None, 0, C.byref(n))
        f = np.zeros((max(n.value, 1), 3))
        self._call("dsph_download_fluid_solid_force", _p(f), n.value, C.byref(n))
        return f[: n.value]

    # ---- frames (binary appended VTU, include/dsph.h)
    def frame_begin(self, slot, i_step, time, species=FRAME_ALL):
        """Packs the species of the current state into frame slot `slot` and
        queues their copy to the host; returns at once (dsph_frame_begin)."""
        self._call("dsph_frame_begin", slot, i_step, time, species)

    def frame_write(self, slot, directory, frame):
        """Writes the frame of `slot` (fluid_NNNN.vtu, solid_NNNN.vtu,
        fluid2Solid_NNNN.vtu, the species it was begun with) into
        `directory` (dsph_frame_write; may run on another thread)."""
        rc = self.L.dsph_frame_write(self.h, slot, os.fsencode(directory), frame)
        if rc != 0:
            raise DsphError("dsph_frame_write", rc, "frame slot empty" if rc == -4 else "file not written")

    def contacts(self):
        n = C.c_int64(0)
        self._call("dsph_download_contacts", None, None, None, None, None, 0, C.byref(n))
        m = n.value
        keys = ("force", "torque", "sliding", "rolling", "torsion")
        out = {k: np.zeros((max(m, 1), 3)) for k in keys}
        self._call("dsph_download_contacts", *[_p(out[k]) for k in keys], m, C.byref(n))
        return {k: v[:m] for k, v in out.items()}

    def bonds(self):
        n = C.c_int64(0)
        self._call("dsph_download_bonds", None, None, None, None, None, 0, C.byref(n))
        m = n.value
        fn, tt = np.zeros(max(m, 1)), np.zeros(max(m, 1))
        fs, tb = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3))
        on = np.zeros(max(m, 1), np.int32)
        self._call("dsph_download_bonds", _p(fn), _p(tt), _p(fs), _p(tb), _p(on, _ip), m, C.byref(n))
        return dict(normal_force=fn[:m], torsion=tt[:m], shear=fs[:m], bending=tb[:m],
                    is_bonded=on[:m])

    def grid_info(self):
        gs = np.zeros(3, np.int32)
        cs = np.zeros(3)
        nc = C.c_int32(0)
        self._call("dsph_grid_info", _p(gs, _ip), _p(cs), C.byref(nc))
        return gs, cs, nc.value

    def grid(self, species):
        ncell = int(self.batch_info()["cell"][-1])  # cells of all scenes
        n = self.nf if species == 0 else self.ns
        h, p = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        s, e = np.zeros(ncell, np.int32), np.zeros(ncell, np.int32)
        self._call("dsph_download_grid", species, _p(h, _ip), _p(p, _ip), _p(s, _ip), _p(e, _ip))
        return dict(hash=h[:n], perm=p[:n], start=s, end=e)

    def status(self):
        w = np.zeros(STATUS_WORDS, np.int64)
        self._call("dsph_status", _p(w, _i64p))
        return w

    def superset_stats(self):
        """Candidates the pair loops walk (rows of the last search):
        {"particles", "walked", "stored", "longest"}."""
        w = np.zeros(4, np.int64)
        self._call("dsph_superset_stats", _p(w, _i64p))
        return dict(particles=int(w[0]), walked=int(w[1]), stored=int(w[2]), longest=int(w[3]))

    def set_scene_time_steps(self, dts):
        """Per-scene time steps of a batch (dsph_set_scene_time_steps); the
        step calls then take dt = 0.  None clears them."""
        if dts is None:
            self._call("dsph_set_scene_time_steps", None)
            return
        d = np.ascontiguousarray(dts, np.float64)
        self._keep_sdt = d
        self._call("dsph_set_scene_time_steps", _p(d))

    def tile_stats(self):
        """The pair-loop tiles of the last superset build: {"tiles",
        "staged", "records", "largest"} (dsph_tile_stats)."""
        w = np.zeros(4, np.int64)
        self._call("dsph_tile_stats", _p(w, _i64p))
        return dict(tiles=int(w[0]), staged=int(w[1]), records=int(w[2]), largest=int(w[3]))

    def hash_sort(self, pos_ptr, n, origin, size, cells, hash_ptr, perm_ptr, start_ptr, end_ptr,
                  noalloc=False):
        """dsph_hash_sort (or dsph_hash_sort_noalloc) on caller device arrays
        (raw device addresses)."""
        o, z = np.asarray(origin, float), np.asarray(size, float)
        g = np.asarray(cells, np.int32)
        self._call("dsph_hash_sort_noalloc" if noalloc else "dsph_hash_sort", _vp(pos_ptr), int(n), _p(o),
                   _p(z), _p(g, _ip), _vp(hash_ptr), _vp(perm_ptr), _vp(start_ptr), _vp(end_ptr))

    def set_timing(self, on=True):
        self._call("dsph_set_timing", int(on))

    def kernel_times(self):
        """{family: (total_ms, launches)} since the last call / set_timing."""
        names = C.create_string_buffer(4096)
        ms = np.zeros(64)
        m = self._call("dsph_kernel_times", names, 4096, _p(ms), 32)
        labels = names.raw.split(b"\0")[:m]
        return {labels[k].decode(): (ms[2 * k], int(ms[2 * k + 1])) for k in range(m)}


# --------------------------------------------------------------- autotuner
TUNING_SPACE = {
    "skin": [0.1, 0.2, 0.3, 0.5],               # Verlet skin / h: superset size vs rebuilds
    "pair_block": [64, 128, 256, 512],          # workgroup of the fluid pair loops
    "build_block": [32, 48, 64],                # particles per wavefront in the superset build
    "group": [1, 2, 4, 8],                      # particles sharing one superset row
    "tile": [0, 1],                             # mixed pair loops over LDS tiles
}


def autotune(scene, precision=PREC_MIXED, flags=0, trial_steps=20, warmup=5, space=None,
             log=None, advance=0):
    """Pick the fastest launch shapes and Verlet skin for a workload.

    Runs the workload (a Scene, or a list of Scenes uploaded as one batch) on
    a throwaway context and times `trial_steps` graph-replayed steps per
    candidate, one coordinate of the search space at a time (LDS tiles, row
    group, skin, pair-loop workgroup, build block), keeping the best value of each
    before moving to the next.  With `advance`, the context first runs that
    many steps and every trial starts from a checkpoint of that state, so
    all candidates time the same window of the moving flow; otherwise the
    trials follow each other from step 1.  Every candidate gives the same
    results; only the time differs.  Returns (config, trials) where config
    feeds Solver.apply_tuning and trials lists (config, ms per step)."""
    space = dict(TUNING_SPACE, **(space or {}))
    g = Solver(precision)
    g.set_flags(flags)
    if isinstance(scene, (list, tuple)):
        g.upload_batch(scene)
        g.set_scene_time_steps([sc.dt for sc in scene])  # each scene at its own step
        dt = 0.0
    else:
        g.upload(scene)
        dt = scene.dt
    g.initialize()
    cfg = {"skin": 0.3, "pair_block": 256, "build_block": 64, "group": 1, "tile": 1}
    trials = []
    step = [1]
    blob = None
    if advance > 0:
        g.run(advance, dt, first=1, graphs=True)
        blob = g.save_state()
        step[0] = advance + 1

    def timed(n):
        g.synchronize()
        t0 = time.perf_counter()
        g.run(n, dt, first=step[0], graphs=True)
        g.synchronize()
        step[0] += n
        return (time.perf_counter() - t0) / n * 1e3

    def measure(c):
        if blob is not None:
            g.load_state(blob)
            step[0] = advance + 1
        g.apply_tuning(c)
        # eager warm-up steps grow any list the new skin overflows; then
        # graph-replayed ones capture the step graphs (every combination of
        # density re-initialisation and record buffer parity), so that no
        # capture falls into the timed steps
        g.run(warmup, dt, first=step[0], graphs=False)
        step[0] += warmup
        g.run(10, dt, first=step[0], graphs=True)
        step[0] += 10
        ms = timed(trial_steps)
        trials.append((dict(c), ms))
        if log:
            log(f"autotune {c} -> {ms:.4f} ms/step")
        return ms

    for key in ("tile", "group", "skin", "pair_block", "build_block"):
        best = None
        for v in space[key]:
            ms = measure(dict(cfg, **{key: v}))
            if best is None or ms < best[0]:
                best = (ms, v)
        cfg[key] = best[1]
    g.close()
    return cfg, trials
