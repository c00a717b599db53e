"""The workload configurations of the benchmark plan as test inputs, and the
compact digests their golden fixtures hold (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from the FP64 CPU oracle).

  cfg0  plain dam break, dx = 0.02: 8,000 fluid + 32,256 wall, 2000 steps
  cfg1  coupled dam break, dx = 0.005: 512,000 fluid + 452,376 wall +
        2,000 Hertz-Mindlin spheres, 1000 steps
  cfg2  granular column: 200,000 spheres (110k free, 12,500 tetrahedral
        clumps of 4, a 40,000-sphere bonded block), no fluid, 5000 steps
  cfg3  precision study, dx = 0.0016: 15,625,000 fluid + 4,281,966 wall +
        20,000 spheres (GPU FP64 against GPU mixed; no oracle fixture)
  cfg4  256-scene dam-break ensemble, artificial viscosity on, one batch,
        each scene stepping with its own time step; the oracle runs a
        seeded subset of its scenes alone

A fixture holds, per configuration:
  - sha256 digests of the four neighbour / contact lists and of the fluid
    and solid hashed grids after initialisation (bit-exact comparisons);
  - the fields after step 1 at a fixed, evenly spread sample of particles;
  - whole-scene integrals at checkpoint steps.
"""
from __future__ import annotations

import hashlib
import math
import os

import numpy as np

import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SAMPLE = 4096          # particles sampled per species for the step-1 fields
FLAG_ART_VISCOSITY = 4  # include/dsph.h

CFG4_SUBSET = (0, 37, 128, 255)   # ensemble scenes the oracle runs alone


def cfg3_scene():
    return scenes.coupled_dam_break(dx=0.0016, n_dem=(50, 20, 20), dem_spacing=0.022,
                                    dem_origin=(0.5, 0.09, 0.012), jitter=0.05, config_id=3,
                                    name="coupled_dam_break_16m")


def ensemble_dts(scs):
    """The time step of every scene of the ensemble: its own."""
    return [sc.dt for sc in scs]


CONFIGS = {
    "cfg0": dict(build=lambda: scenes.dam_break(0.02), steps=2000,
                 checkpoints=(1, 10, 100, 500, 1000, 2000), flags=0),
    "cfg1": dict(build=lambda: scenes.coupled_dam_break(dx=0.005), steps=1000,
                 checkpoints=(1, 10, 100, 250, 500, 1000), flags=0),
    "cfg2": dict(build=lambda: scenes.granular_column(n_sphere=110000, n_clump=12500, n_bonded=40000),
                 steps=5000, checkpoints=(1, 10, 100, 1000, 5000), flags=0),
}
CFG4_STEPS = 1000
CFG4_CHECKPOINTS = (1, 100, 1000)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(np.asarray(a.shape, np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def list_digests(x, kinds=(0, 1, 2, 3)):
    """{kind: (n_pairs, digest of pointed, pointing and counts)} of a
    solver or oracle (same read-back API)."""
    out = {}
    for kind in kinds:
        a, b, c = x.list(kind)
        a, b = np.asarray(a, np.int32), np.asarray(b, np.int32)
        c = np.asarray(c, np.int32) if kind != 3 else np.zeros(0, np.int32)
        out[kind] = (len(a), digest(a, b, c))
    return out


def grid_digests(x):
    out = {}
    for species in (0, 1):
        g = x.grid(species)
        out[species] = digest(*(np.asarray(g[k], np.int32) for k in ("hash", "perm", "start", "end")))
    return out


def sample_index(n, m=SAMPLE):
    if n <= m:
        return np.arange(n)
    return np.unique(np.linspace(0, n - 1, m).astype(np.int64))


def moving(sc):
    return np.flatnonzero(np.asarray(sc.solid_inv_mass) > 0)


def step1_fields(x, sc):
    """The step-1 state at the sample: fluid position, velocity, acceleration
    and dRho; moving solids' position, velocity, acceleration, angular
    velocity and torque."""
    out = {}
    if sc.n_fluid:
        f = x.fluid()
        i = sample_index(sc.n_fluid)
        for k in ("pos", "vel", "acc", "drho"):
            out["f_" + k] = f[k][i]
    mov = moving(sc)
    if len(mov):
        s = x.solid()
        j = mov[sample_index(len(mov))]
        for k in ("pos", "vel", "acc", "angvel", "torque"):
            out["s_" + k] = s[k][j]
    return out


def integrals(x, sc, fluid_range=None, solid_range=None):
    """Whole-scene integrals: fluid volume, kinetic and potential energy,
    surge front and mean density; the moving solids' centroid, kinetic
    energy and (with bonds) the intact bond count.  With ranges, those of
    one scene of a batch (caller indices [a, b))."""
    out = {}
    if sc.n_fluid:
        f = x.fluid()
        a, b = fluid_range if fluid_range else (0, sc.n_fluid)
        m = (np.asarray(sc.fluid_h) / 1.3) ** 3 * np.asarray(sc.fluid_rho0)
        rho = np.asarray(sc.fluid_rho0) + f["drho"][a:b]
        v, p = f["vel"][a:b], f["pos"][a:b]
        out.update(volume=float(np.sum(m / rho)), ke=float(0.5 * np.sum(m * np.sum(v * v, axis=1))),
                   pe=float(np.sum(m * 9.81 * p[:, 2])), front=float(np.max(p[:, 0])),
                   mean_rho=float(np.mean(rho)))
    mov = moving(sc)
    if len(mov):
        s = x.solid()
        a = solid_range[0] if solid_range else 0
        pos, vel = s["pos"][a + mov], s["vel"][a + mov]
        ms = 1.0 / np.asarray(sc.solid_inv_mass)[mov]
        out.update(dem_x=float(np.mean(pos[:, 0])), dem_y=float(np.mean(pos[:, 1])),
                   dem_z=float(np.mean(pos[:, 2])),
                   dem_ke=float(0.5 * np.sum(ms * np.sum(vel * vel, axis=1))),
                   dem_pe=float(np.sum(ms * 9.81 * pos[:, 2])))
        if solid_range is None and np.any(np.asarray(sc.solid_cluster) >= 0):
            out["bonds_intact"] = float(np.sum(x.bonds()["is_bonded"]))
    return out


def integral_floor(key, ref):
    """Scale below which a relative error is not meaningful: kinetic
    energies start at zero and are compared against 1e-6 of the potential
    energy (about the kinetic energy after 1 ms of free fall); a centroid
    coordinate against 1 mm."""
    if key == "ke":
        return 1e-6 * abs(ref.get("pe", 0.0)) or 1e-12
    if key == "dem_ke":
        return 1e-6 * abs(ref.get("dem_pe", 0.0)) or 1e-12
    if key in ("dem_x", "dem_y", "dem_z", "front"):
        return 1e-3
    return 1e-12


def integral_deviations(got, ref):
    return {k: abs(got[k] - ref[k]) / max(abs(ref[k]), integral_floor(k, ref)) for k in ref}


def fixture_path(name):
    return os.path.join(GOLDEN, f"{name}.npz")


def save(name, d):
    flat = {}
    for k, v in d.items():
        flat[k] = np.asarray(v)
    np.savez_compressed(fixture_path(name), **flat)


def load(name):
    with np.load(fixture_path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def pack(lists, grids, fields, ints, checkpoints, meta):
    """Fixture dictionary from the pieces (all plain arrays and strings)."""
    d = dict(meta)
    for kind, (n, h) in lists.items():
        d[f"list{kind}_n"] = n
        d[f"list{kind}_sha"] = h
    for sp, h in grids.items():
        d[f"grid{sp}_sha"] = h
    for k, v in fields.items():
        d["step1_" + k] = v
    d["checkpoints"] = np.asarray(checkpoints, np.int64)
    keys = sorted(ints[checkpoints[0]])
    d["integral_keys"] = np.asarray(keys)
    d["integrals"] = np.asarray([[ints[c][k] for k in keys] for c in checkpoints])
    return d


def unpack_integrals(d):
    keys = [str(k) for k in d["integral_keys"]]
    return {int(c): dict(zip(keys, map(float, row))) for c, row in zip(d["checkpoints"], d["integrals"])}
