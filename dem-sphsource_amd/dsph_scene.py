"""Scene container and its ctypes view (the ``dsph_scene`` struct of include/dsph.h).

A ``Scene`` is the reference's ``HostData`` (core/hostStructs.h:7-399) as a set
of numpy arrays: fluid, solid (walls, spheres, clump pebbles), clumps,
infinite walls and the material-pair contact tables.  ``Scene.cstruct()``
builds the C struct the library and the oracle both accept; the arrays stay
owned by the Scene and must outlive the call that uses the struct.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class CScene(C.Structure):
    _fields_ = [
        ("domain_origin", C.c_double * 3),
        ("domain_size", C.c_double * 3),
        ("gravity", C.c_double * 3),
        ("n_fluid", C.c_int32),
        ("fluid_pos", _dp), ("fluid_vel", _dp), ("fluid_h", _dp),
        ("fluid_rho0", _dp), ("fluid_drho", _dp), ("fluid_c", _dp), ("fluid_nu", _dp),
        ("n_solid", C.c_int32),
        ("solid_pos", _dp), ("solid_vel", _dp), ("solid_angvel", _dp), ("solid_normal", _dp),
        ("solid_radius", _dp), ("solid_eff_radius", _dp), ("solid_inv_mass", _dp),
        ("solid_material", _ip), ("solid_cluster", _ip), ("solid_clump", _ip),
        ("solid_dummy", _ip),
        ("n_clump", C.c_int32),
        ("clump_centroid", _dp), ("clump_vel", _dp), ("clump_angvel", _dp),
        ("clump_orientation", _dp), ("clump_inv_inertia", _dp), ("clump_inv_mass", _dp),
        ("clump_pebble_start", _ip), ("clump_pebble_end", _ip),
        ("n_wall", C.c_int32),
        ("wall_point", _dp), ("wall_normal", _dp), ("wall_material", _ip),
        ("n_material", C.c_int32),
        ("hertz", _dp), ("linear", _dp), ("bonded", _dp),
        ("fluid_alpha", _dp),
        ("n_cylinder", C.c_int32),
        ("cylinder_point", _dp), ("cylinder_axis", _dp), ("cylinder_angvel", _dp),
        ("cylinder_radius", _dp), ("cylinder_material", _ip),
    ]


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def n_comb(n_material: int) -> int:
    return n_material * (n_material + 1) // 2


def combined_index(n_material: int, a: int, b: int) -> int:
    """Material-pair table index (core/deviceStructs.h:958-973)."""
    if a >= n_material or b >= n_material:
        return -1
    i, j = (a, b) if a <= b else (b, a)
    return (i * (2 * n_material - i + 1)) // 2 + j - i


@dataclass
class Scene:
    domain_origin: np.ndarray
    domain_size: np.ndarray
    gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, -9.81]))
    # fluid
    fluid_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    fluid_vel: np.ndarray | None = None
    fluid_h: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fluid_rho0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fluid_drho: np.ndarray | None = None
    fluid_c: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fluid_nu: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # solid
    solid_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    solid_vel: np.ndarray | None = None
    solid_angvel: np.ndarray | None = None
    solid_normal: np.ndarray | None = None
    solid_radius: np.ndarray = field(default_factory=lambda: np.zeros(0))
    solid_eff_radius: np.ndarray = field(default_factory=lambda: np.zeros(0))
    solid_inv_mass: np.ndarray = field(default_factory=lambda: np.zeros(0))
    solid_material: np.ndarray | None = None
    solid_cluster: np.ndarray | None = None
    solid_clump: np.ndarray | None = None
    solid_dummy: np.ndarray | None = None
    # clumps
    clump_centroid: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    clump_vel: np.ndarray | None = None
    clump_angvel: np.ndarray | None = None
    clump_orientation: np.ndarray | None = None
    clump_inv_inertia: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    clump_inv_mass: np.ndarray = field(default_factory=lambda: np.zeros(0))
    clump_pebble_start: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    clump_pebble_end: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    # infinite walls
    wall_point: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    wall_normal: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    wall_material: np.ndarray | None = None
    # contact tables (field-major, see include/dsph.h)
    n_material: int = 10
    hertz: np.ndarray | None = None
    linear: np.ndarray | None = None
    bonded: np.ndarray | None = None
    fluid_alpha: np.ndarray | None = None  # artificial viscosity per fluid particle
    # cylinder walls
    cylinder_point: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    cylinder_axis: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    cylinder_angvel: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    cylinder_radius: np.ndarray = field(default_factory=lambda: np.zeros(0))
    cylinder_material: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    # driver parameters (not part of the C struct)
    dt: float = 1e-4
    gap: int = 1
    name: str = ""

    def __post_init__(self):
        nc = n_comb(self.n_material)
        if self.hertz is None:
            # HostHertzianContactModel defaults: res = 1 (hostStructs.h:305)
            self.hertz = np.zeros((8, nc))
            self.hertz[2, :] = 1.0
        if self.linear is None:
            self.linear = np.zeros((11, nc))
        if self.bonded is None:
            # HostBondedContactModel defaults: gamma 1, kn/ks 2 (hostStructs.h:346)
            self.bonded = np.zeros((6, nc))
            self.bonded[0, :] = 1.0
            self.bonded[2, :] = 2.0

    # ------------------------------------------------------------------
    @property
    def n_fluid(self) -> int:
        return int(np.asarray(self.fluid_pos).reshape(-1, 3).shape[0])

    @property
    def n_solid(self) -> int:
        return int(np.asarray(self.solid_pos).reshape(-1, 3).shape[0])

    @property
    def n_clump(self) -> int:
        return int(np.asarray(self.clump_centroid).reshape(-1, 3).shape[0])

    @property
    def n_wall(self) -> int:
        return int(np.asarray(self.wall_point).reshape(-1, 3).shape[0])

    @property
    def n_cylinder(self) -> int:
        return int(np.asarray(self.cylinder_point).reshape(-1, 3).shape[0])

    @property
    def n_particles(self) -> int:
        return self.n_fluid + self.n_solid

    def set_hertz(self, a, b, E, G, res, kr_ks, kt_ks, mu_s, mu_r, mu_t):
        """dataBase::setHertzianContactModel (core/dataBase.h:111-128)."""
        c = combined_index(self.n_material, a, b)
        self.hertz[:, c] = [E, G, res, kr_ks, kt_ks, mu_s, mu_r, mu_t]

    def set_linear(self, a, b, *vals):
        """dataBase::setLinearContactModel (core/dataBase.h:130-150)."""
        c = combined_index(self.n_material, a, b)
        self.linear[:, c] = vals

    def set_bonded(self, a, b, E, kn_ks, gamma, sigma_s, C_, mu):
        """dataBase::setBondedContactModel (core/dataBase.h:152-167)."""
        c = combined_index(self.n_material, a, b)
        self.bonded[:, c] = [gamma, E, kn_ks, sigma_s, C_, mu]

    # ------------------------------------------------------------------
    def cstruct(self) -> CScene:
        keep = []

        def dp(a, width=None):
            if a is None:
                return None
            arr = _f64(a).reshape(-1)
            keep.append(arr)
            return arr.ctypes.data_as(_dp)

        def ip(a):
            if a is None:
                return None
            arr = _i32(a).reshape(-1)
            keep.append(arr)
            return arr.ctypes.data_as(_ip)

        s = CScene()
        s.domain_origin[:] = [float(x) for x in self.domain_origin]
        s.domain_size[:] = [float(x) for x in self.domain_size]
        s.gravity[:] = [float(x) for x in self.gravity]
        s.n_fluid = self.n_fluid
        s.fluid_pos = dp(self.fluid_pos)
        s.fluid_vel = dp(self.fluid_vel)
        s.fluid_h = dp(self.fluid_h)
        s.fluid_rho0 = dp(self.fluid_rho0)
        s.fluid_drho = dp(self.fluid_drho)
        s.fluid_c = dp(self.fluid_c)
        s.fluid_nu = dp(self.fluid_nu)
        s.n_solid = self.n_solid
        s.solid_pos = dp(self.solid_pos)
        s.solid_vel = dp(self.solid_vel)
        s.solid_angvel = dp(self.solid_angvel)
        s.solid_normal = dp(self.solid_normal)
        s.solid_radius = dp(self.solid_radius)
        s.solid_eff_radius = dp(self.solid_eff_radius)
        s.solid_inv_mass = dp(self.solid_inv_mass)
        s.solid_material = ip(self.solid_material)
        s.solid_cluster = ip(self.solid_cluster)
        s.solid_clump = ip(self.solid_clump)
        s.solid_dummy = ip(self.solid_dummy)
        s.n_clump = self.n_clump
        s.clump_centroid = dp(self.clump_centroid)
        s.clump_vel = dp(self.clump_vel)
        s.clump_angvel = dp(self.clump_angvel)
        s.clump_orientation = dp(self.clump_orientation)
        s.clump_inv_inertia = dp(self.clump_inv_inertia)
        s.clump_inv_mass = dp(self.clump_inv_mass)
        s.clump_pebble_start = ip(self.clump_pebble_start)
        s.clump_pebble_end = ip(self.clump_pebble_end)
        s.n_wall = self.n_wall
        s.wall_point = dp(self.wall_point)
        s.wall_normal = dp(self.wall_normal)
        s.wall_material = ip(self.wall_material)
        s.n_material = int(self.n_material)
        s.hertz = dp(self.hertz)
        s.linear = dp(self.linear)
        s.bonded = dp(self.bonded)
        s.fluid_alpha = dp(self.fluid_alpha)
        s.n_cylinder = self.n_cylinder
        s.cylinder_point = dp(self.cylinder_point)
        s.cylinder_axis = dp(self.cylinder_axis)
        s.cylinder_angvel = dp(self.cylinder_angvel)
        s.cylinder_radius = dp(self.cylinder_radius)
        s.cylinder_material = ip(self.cylinder_material)
        s._keep = keep  # keep arrays alive with the struct
        return s
