"""Deterministic scene builders for the dam-break / DEM workloads.

Every builder reproduces the reference's own scene-building arithmetic:
``getRegularPackedPoints`` (core/generatePoints.h:5-34) for lattices, and the
``addFluid`` / ``addDummyParticleWall`` / ``addSolid`` / ``addClump`` /
``addInfiniteWall`` field rules of core/dataBase.cpp:3-177 (effective radii
1.3dx for fluid, 2.6r for wall dummies, 1.1r for DEM spheres; sphere inverse
mass 3/(4 pi r^3 rho); pebble inverse mass 1/((2r)^3 rho_clump)).  Random
draws come from a counter-based SplitMix64 stream so the CPU oracle and the
GPU receive bit-identical inputs.

Workloads (sizes as in the project's benchmark plan):
  dam_break(dx=0.02)            plain 3-D dam break, ~8k fluid + 32,256 wall
  coupled_dam_break(dx=0.005)   ~512k fluid + 452,376 wall + 2,000 DEM spheres
  granular_column(...)          spheres, tetrahedral clumps and a bonded block
  random_* helpers              small randomised scenes for parity tests
"""
from __future__ import annotations

import math

import numpy as np

from dsph_scene import Scene, combined_index

ROOT_SEED = 0x20261203
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


# --------------------------------------------------------------------------
# random numbers
# --------------------------------------------------------------------------
def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    """n outputs of a counter-based SplitMix64 (Steele, Lea & Flood 2014)."""
    base = np.uint64((seed ^ ((stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF))
                     & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = base + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int, lo=0.0, hi=1.0) -> np.ndarray:
    u = (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


# --------------------------------------------------------------------------
# lattice: getRegularPackedPoints (core/generatePoints.h:5-34)
# --------------------------------------------------------------------------
def regular_packed_points(origin, size, spacing) -> np.ndarray:
    origin = [float(v) for v in origin]
    size = [float(v) for v in size]
    num = [int((s + 1.e-10) / spacing) for s in size]
    if min(num) == 0:
        return np.array([[origin[k] + 0.5 * size[k] for k in range(3)]])
    axes = []
    for k in range(3):
        d = size[k] / float(num[k])
        axes.append(origin[k] + d * (np.arange(1, num[k] + 1, dtype=np.float64) - 0.5))
    X, Y, Z = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")  # x outer, z inner
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


# --------------------------------------------------------------------------
# builders
# --------------------------------------------------------------------------
def _append_solid(sc: Scene, pos, radius, eff, invm, material=0, cluster=-1, clump=-1,
                  dummy=0, vel=None, angvel=None):
    n = len(pos)

    def cat(old, new, dtype=np.float64, width=None):
        new = np.asarray(new, dtype=dtype)
        new = np.broadcast_to(new, (n, width) if width else (n,))
        if old is None:
            return np.array(new)
        return np.concatenate([np.asarray(old, dtype=dtype).reshape((-1, width) if width else -1), new])

    sc.solid_pos = cat(sc.solid_pos, pos, width=3)
    sc.solid_vel = cat(sc.solid_vel, vel if vel is not None else np.zeros((n, 3)), width=3)
    sc.solid_angvel = cat(sc.solid_angvel, angvel if angvel is not None else np.zeros((n, 3)), width=3)
    sc.solid_radius = cat(sc.solid_radius, radius)
    sc.solid_eff_radius = cat(sc.solid_eff_radius, eff)
    sc.solid_inv_mass = cat(sc.solid_inv_mass, invm)
    sc.solid_material = cat(sc.solid_material, material, np.int32)
    sc.solid_cluster = cat(sc.solid_cluster, cluster, np.int32)
    sc.solid_clump = cat(sc.solid_clump, clump, np.int32)
    sc.solid_dummy = cat(sc.solid_dummy, dummy, np.int32)


def empty_scene(origin, size, n_material=10, name="") -> Scene:
    sc = Scene(domain_origin=np.asarray(origin, float), domain_size=np.asarray(size, float),
               n_material=n_material, name=name)
    sc.solid_pos = np.zeros((0, 3))
    sc.solid_vel = np.zeros((0, 3))
    sc.solid_angvel = np.zeros((0, 3))
    sc.solid_radius = np.zeros(0)
    sc.solid_eff_radius = np.zeros(0)
    sc.solid_inv_mass = np.zeros(0)
    sc.solid_material = np.zeros(0, np.int32)
    sc.solid_cluster = np.zeros(0, np.int32)
    sc.solid_clump = np.zeros(0, np.int32)
    sc.solid_dummy = np.zeros(0, np.int32)
    return sc


def add_fluid(sc: Scene, pos, h, rho0, c, nu, vel=(0.0, 0.0, 0.0)):
    """dataBase::addFluid (core/dataBase.cpp:3-17)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    n = len(pos)
    cat = lambda old, new: np.concatenate([np.asarray(old, float).reshape(-1), new])
    old_n = sc.n_fluid
    sc.fluid_pos = np.concatenate([np.asarray(sc.fluid_pos, float).reshape(-1, 3), pos])
    v = np.broadcast_to(np.asarray(vel, float), (n, 3))
    sc.fluid_vel = np.concatenate([np.zeros((old_n, 3)) if sc.fluid_vel is None else sc.fluid_vel, v])
    sc.fluid_h = cat(sc.fluid_h, np.full(n, h))
    sc.fluid_rho0 = cat(sc.fluid_rho0, np.full(n, rho0))
    sc.fluid_c = cat(sc.fluid_c, np.full(n, c))
    sc.fluid_nu = cat(sc.fluid_nu, np.full(n, nu))


def add_dummy_wall(sc: Scene, pos, radius):
    """dataBase::addDummyParticleWall (core/dataBase.cpp:19-31)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    _append_solid(sc, pos, radius, 2.6 * radius, 0.0, dummy=1)


def sphere_inv_mass(radius, density):
    """1 / (4/3 r^3 pi rho) as evaluated in core/dataBase.cpp:49."""
    radius = np.asarray(radius, float)
    return np.where(density > 0, 1.0 / (4.0 / 3.0 * np.power(radius, 3) * math.pi * density), 0.0)


def add_solid(sc: Scene, pos, radius, density, material=0, vel=(0, 0, 0)):
    """dataBase::addSolid (core/dataBase.cpp:33-53)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    n = len(pos)
    radius = np.broadcast_to(np.asarray(radius, float), (n,))
    _append_solid(sc, pos, radius, 1.1 * radius, sphere_inv_mass(radius, density),
                  material=material, vel=np.broadcast_to(np.asarray(vel, float), (n, 3)))


def add_cluster(sc: Scene, pos, radius, density, material=0, vel=(0, 0, 0)):
    """dataBase::addCluster before the run (core/dataBase.cpp:84-118)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    n = len(pos)
    cid = 0
    if sc.n_solid > 0:
        cid = int(np.max(sc.solid_cluster)) + 1
    radius = np.broadcast_to(np.asarray(radius, float), (n,))
    density = np.broadcast_to(np.asarray(density, float), (n,))
    _append_solid(sc, pos, radius, 1.1 * radius, sphere_inv_mass(radius, density),
                  material=material, cluster=cid,
                  vel=np.broadcast_to(np.asarray(vel, float), (n, 3)))


def inverse_sym(I):
    """symMatrix inverse (core/myMat.h:129-142); I = (xx,yy,zz,xy,xz,yz)."""
    xx, yy, zz, xy, xz, yz = I
    det = 1.0 / (xx * (yy * zz - yz * yz) + xy * (yz * xz - xy * zz) + xz * (xy * yz - yy * xz))
    return np.array([(yy * zz - yz * yz) * det, (xx * zz - xz * xz) * det,
                     (xx * yy - xy * xy) * det, -(xy * zz - xz * yz) * det,
                     (xy * yz - xz * yy) * det, -(xx * yz - xz * xy) * det])


def add_clump(sc: Scene, pos, radius, centroid, mass, inertia, material=0,
              vel=(0, 0, 0), angvel=(0, 0, 0)):
    """dataBase::addClump before the run (core/dataBase.cpp:120-177)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    radius = np.asarray(radius, float)
    centroid = np.asarray(centroid, float)
    vel = np.asarray(vel, float)
    angvel = np.asarray(angvel, float)
    start = sc.n_solid
    # summed in pebble order, as the reference's loop does (dataBase.cpp:138-142)
    volume = 0.0
    for v in np.power(2.0 * radius, 3):
        volume += float(v)
    dens = mass / volume if volume > 0 else 0.0
    invm = np.where((dens > 0) & (radius > 0), 1.0 / (np.power(2.0 * radius, 3) * dens), 0.0)
    pv = vel + np.cross(angvel, pos - centroid)
    _append_solid(sc, pos, radius, 1.1 * radius, invm, material=material,
                  clump=sc.n_clump, vel=pv, angvel=np.broadcast_to(angvel, (len(pos), 3)))
    end = sc.n_solid
    sc.clump_centroid = np.concatenate([sc.clump_centroid.reshape(-1, 3), centroid[None]])
    sc.clump_vel = np.concatenate([np.zeros((0, 3)) if sc.clump_vel is None else sc.clump_vel, vel[None]])
    sc.clump_angvel = np.concatenate([np.zeros((0, 3)) if sc.clump_angvel is None else sc.clump_angvel,
                                      angvel[None]])
    sc.clump_inv_inertia = np.concatenate([sc.clump_inv_inertia.reshape(-1, 6),
                                           inverse_sym(inertia)[None]])
    sc.clump_inv_mass = np.concatenate([sc.clump_inv_mass, [1.0 / mass if mass > 0 else 0.0]])
    sc.clump_pebble_start = np.concatenate([sc.clump_pebble_start, [start]]).astype(np.int32)
    sc.clump_pebble_end = np.concatenate([sc.clump_pebble_end, [end]]).astype(np.int32)


def add_infinite_wall(sc: Scene, point, normal, material=0):
    """dataBase::addInfiniteWall (core/dataBase.cpp:55-67)."""
    sc.wall_point = np.concatenate([sc.wall_point.reshape(-1, 3), np.asarray(point, float)[None]])
    sc.wall_normal = np.concatenate([sc.wall_normal.reshape(-1, 3), np.asarray(normal, float)[None]])
    wm = np.zeros(0, np.int32) if sc.wall_material is None else sc.wall_material
    sc.wall_material = np.concatenate([wm, [material]]).astype(np.int32)


def add_cylinder_wall(sc: Scene, point, axis, angvel, radius, material=0):
    """dataBase::addCylinderWall (core/dataBase.cpp:69-82), with the radius
    kept (the reference drops it)."""
    sc.cylinder_point = np.concatenate([sc.cylinder_point.reshape(-1, 3), np.asarray(point, float)[None]])
    sc.cylinder_axis = np.concatenate([sc.cylinder_axis.reshape(-1, 3), np.asarray(axis, float)[None]])
    sc.cylinder_angvel = np.concatenate([sc.cylinder_angvel.reshape(-1, 3), np.asarray(angvel, float)[None]])
    sc.cylinder_radius = np.concatenate([sc.cylinder_radius, [float(radius)]])
    sc.cylinder_material = np.concatenate([sc.cylinder_material, [material]]).astype(np.int32)


# --------------------------------------------------------------------------
# the workloads
# --------------------------------------------------------------------------
G = 9.81
RHO0 = 1000.0
NU = 1.0e-6


def sound_speed(column_height=0.4):
    return 10.0 * math.sqrt(2.0 * G * column_height)


def tank_wall_points(origin, size, spacing, L, W) -> np.ndarray:
    """The points of regular_packed_points(origin, size, spacing) outside
    [0,L]x[0,W]x[0,inf) (the wall layers of an open tank), in the same
    order, without materialising the whole lattice: the lattice's y-z slab
    is filtered once and reused for every x inside [0, L]."""
    origin = [float(v) for v in origin]
    size = [float(v) for v in size]
    num = [int((s + 1.e-10) / spacing) for s in size]
    if min(num) == 0:
        p = regular_packed_points(origin, size, spacing)
        keep = (p[:, 0] < 0) | (p[:, 1] < 0) | (p[:, 2] < 0) | (p[:, 0] > L) | (p[:, 1] > W)
        return p[keep]
    ax = [origin[k] + (size[k] / float(num[k])) * (np.arange(1, num[k] + 1, dtype=np.float64) - 0.5)
          for k in range(3)]
    Y, Z = np.meshgrid(ax[1], ax[2], indexing="ij")
    yz = np.stack([Y.ravel(), Z.ravel()], axis=1)
    shell = yz[(yz[:, 0] < 0) | (yz[:, 1] < 0) | (yz[:, 0] > W)]
    parts = []
    for x in ax[0]:
        q = yz if (x < 0 or x > L) else shell
        parts.append(np.concatenate([np.full((len(q), 1), x), q], axis=1))
    return np.concatenate(parts) if parts else np.zeros((0, 3))


def dam_break(dx=0.02, L=1.6, W=0.6, H=0.6, column=(0.4, 0.4, 0.4), with_floor_walls=False,
              name="dam_break") -> Scene:
    """Plain 3-D dam break (damBreak.cpp:14-58 pattern): tank LxWxH, water
    column in the corner, 3 lattice layers of wall dummies outside
    [0,L]x[0,W]x[0,inf) (open top), domain origin -3dx and size +6dx."""
    o = np.array([-3 * dx] * 3)
    size = np.array([L + 6 * dx, W + 6 * dx, H + 6 * dx])
    sc = empty_scene(o, size, name=name)
    h = 1.3 * dx
    c0 = sound_speed(column[2])
    fp = regular_packed_points((0, 0, 0), column, dx)
    add_fluid(sc, fp, h, RHO0, c0, NU)
    add_dummy_wall(sc, tank_wall_points(o, size, dx, L, W), 0.5 * dx)
    sc.dt = 0.25 * h / c0
    sc.gap = 1
    return sc


def hertz_effective(E, nu_p):
    """Effective moduli fed to 4/3 E* sqrt(R d) and 8 G* sqrt(R d)
    (SPHKernelDEMContactModel.h:289-290): E* = E/(2(1-nu^2)), G* = G/(2(2-nu))."""
    Gm = E / (2 * (1 + nu_p))
    return E / (2 * (1 - nu_p ** 2)), Gm / (2 * (2 - nu_p))


def add_tank_planes(sc: Scene, L, W, material=0):
    """Floor first (the wall loop stops at the first non-touching wall)."""
    add_infinite_wall(sc, (0, 0, 0), (0, 0, 1), material)
    add_infinite_wall(sc, (0, 0, 0), (1, 0, 0), material)
    add_infinite_wall(sc, (L, 0, 0), (-1, 0, 0), material)
    add_infinite_wall(sc, (0, 0, 0), (0, 1, 0), material)
    add_infinite_wall(sc, (0, W, 0), (0, -1, 0), material)


def coupled_dam_break(dx=0.005, n_dem=(20, 10, 10), dem_spacing=0.025,
                      dem_origin=(0.9, 0.175, 0.012), r=0.01, rho_s=2500.0, jitter=0.1,
                      E=1.0e7, nu_p=0.3, res=0.5, mu=0.4, config_id=1,
                      name="coupled_dam_break") -> Scene:
    """Coupled DEM-SPH dam break: the plain dam break at dx plus a seeded
    jittered lattice of DEM spheres downstream, Hertz-Mindlin contact and
    tank planes."""
    sc = dam_break(dx=dx, name=name)
    nx, ny, nz = n_dem
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([dem_origin[0] + dem_spacing * ii.ravel(),
                     dem_origin[1] + dem_spacing * jj.ravel(),
                     dem_origin[2] + dem_spacing * kk.ravel()], axis=1)
    n = len(base)
    jit = uniform(ROOT_SEED, 1000 * config_id + 1, 3 * n, -jitter * r, jitter * r).reshape(n, 3)
    add_solid(sc, base + jit, r, rho_s, material=0)
    Es, Gs = hertz_effective(E, nu_p)
    sc.set_hertz(0, 0, Es, Gs, res, 0.0, 0.0, mu, 0.0, 0.0)
    add_tank_planes(sc, 1.6, 0.6)
    k = 0.5 * E * math.pi * r
    mij = 8 * r ** 3 * rho_s
    dt_dem = 0.2 * math.sqrt(mij / k)
    sc.dt = min(sc.dt, dt_dem)
    sc.gap = 1
    return sc


def tetra_offsets(edge):
    """Vertices of a regular tetrahedron of the given edge, centred at 0."""
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    return v * (edge / (2.0 * math.sqrt(2.0)))


def granular_column(n_sphere=4000, n_clump=250, n_bonded=1000, r0=0.002, rho_s=2500.0,
                    E=1.0e7, nu_p=0.3, res=0.5, mu=0.4, k_bond=1.0e8, sigma_t=1.0e6,
                    config_id=2, name="granular_column") -> Scene:
    """Granular column collapse, no fluid: free spheres with radii
    U[0.8,1.2]*r0, 4-sphere regular-tetrahedron clumps with touching pebbles,
    and a bonded block (parallel bonds) of touching spheres of radius r0,
    released onto the plane z = 0.  Column footprint centred at x = y = 0.5."""
    sc = empty_scene((0.3, 0.3, -0.01), (0.4, 0.4, 1.0), name=name)
    rmax = 1.2 * r0
    # bonded block: cubic lattice of touching spheres, bottom of the column
    side = max(1, int(round(n_bonded ** (1.0 / 3.0))))
    nb_layer = side * side
    idx = np.arange(n_bonded)
    bx, by, bz = idx % side, (idx // side) % side, idx // nb_layer
    s = 2 * r0
    cx = 0.5 - 0.5 * (side - 1) * s
    bpos = np.stack([cx + s * bx, cx + s * by, r0 + s * bz], axis=1).astype(float)
    z_top = float(bpos[:, 2].max() + r0) if n_bonded else 0.0
    if n_bonded:
        add_cluster(sc, bpos, r0, rho_s, material=0)
    # clumps: one per cell of a coarse lattice above the block
    cell = 2.0 * (rmax + 2 * rmax * math.sqrt(6.0) / 4.0) * 1.05
    per = max(1, int(0.2 / cell))
    u = uniform(ROOT_SEED, 1000 * config_id + 2, max(n_clump, 1), 0.8, 1.2)
    for c in range(n_clump):
        ix, iy, iz = c % per, (c // per) % per, c // (per * per)
        cen = np.array([0.5 - 0.5 * (per - 1) * cell + ix * cell,
                        0.5 - 0.5 * (per - 1) * cell + iy * cell,
                        z_top + 0.5 * cell + iz * cell])
        rc = r0 * u[c]
        pts = cen + tetra_offsets(2 * rc)
        m_one = 4.0 / 3.0 * math.pi * rc ** 3 * rho_s
        mass = 4 * m_one
        # inertia of 4 spheres about the centroid (parallel-axis theorem)
        d2 = float(np.sum(tetra_offsets(2 * rc)[0] ** 2))
        Ii = 4 * (0.4 * m_one * rc * rc) + mass * d2 * 2.0 / 3.0
        add_clump(sc, pts, np.full(4, rc), cen, mass, (Ii, Ii, Ii, 0.0, 0.0, 0.0), material=0)
    z_top2 = z_top + (math.ceil(n_clump / (per * per)) * cell if n_clump else 0.0)
    # free spheres: lattice of spacing 2*rmax*1.02 above the clumps
    sp = 2 * rmax * 1.02
    per_s = max(1, int(0.2 / sp))
    k = np.arange(n_sphere)
    sx, sy, sz = k % per_s, (k // per_s) % per_s, k // (per_s * per_s)
    spos = np.stack([0.5 - 0.5 * (per_s - 1) * sp + sx * sp,
                     0.5 - 0.5 * (per_s - 1) * sp + sy * sp,
                     z_top2 + rmax + sz * sp], axis=1)
    rad = r0 * uniform(ROOT_SEED, 1000 * config_id + 3, max(n_sphere, 1), 0.8, 1.2)[:n_sphere]
    add_solid(sc, spos, rad, rho_s, material=0)
    Es, Gs = hertz_effective(E, nu_p)
    sc.set_hertz(0, 0, Es, Gs, res, 0.0, 0.0, mu, 0.0, 0.0)
    # bond stiffness k_n = E_b * pi r^2 / (2r) -> E_b = k_n (rA+rB) / (pi rmin^2)
    Eb = k_bond * (2 * r0) / (math.pi * r0 * r0)
    sc.set_bonded(0, 0, Eb, 1.0, 1.0, sigma_t, 0.0, 0.0)
    add_infinite_wall(sc, (0, 0, 0), (0, 0, 1), 0)
    m = 4.0 / 3.0 * math.pi * (0.8 * r0) ** 3 * rho_s
    sc.dt = 0.2 * 2.0 / math.sqrt(k_bond / m) / 2.0
    sc.gap = 1
    # domain: the column footprint plus room to spread, up to above the top
    max_z = float(np.max(sc.solid_pos[:, 2])) + 0.02
    sc.domain_origin = np.array([0.3, 0.3, -0.01])
    sc.domain_size = np.array([0.4, 0.4, max_z + 0.01])
    return sc


def random_fluid_box(n=2000, box=(0.1, 0.08, 0.06), dx=0.005, seed_stream=7, h_jitter=0.0,
                     with_walls=True, vel_scale=0.1, drho_scale=1.0, name="random_fluid") -> Scene:
    """Jittered fluid block inside a walled box: irregular neighbour counts,
    non-zero velocities and densities (parity test input)."""
    L, W, H = box
    o = np.array([-3 * dx] * 3)
    size = np.array([L + 6 * dx, W + 6 * dx, H + 6 * dx])
    sc = empty_scene(o, size, name=name)
    fp = regular_packed_points((0, 0, 0), (L, W, H * 0.8), dx)
    m = len(fp)
    jit = uniform(ROOT_SEED, seed_stream, 3 * m, -0.3 * dx, 0.3 * dx).reshape(m, 3)
    fp = fp + jit
    h = 1.3 * dx * (1.0 + h_jitter * uniform(ROOT_SEED, seed_stream + 1, m, -1, 1))
    c0 = sound_speed(0.4)
    add_fluid(sc, fp, 1.3 * dx, RHO0, c0, NU)
    sc.fluid_h = h
    sc.fluid_vel = uniform(ROOT_SEED, seed_stream + 2, 3 * m, -vel_scale, vel_scale).reshape(m, 3)
    sc.fluid_drho = uniform(ROOT_SEED, seed_stream + 3, m, -drho_scale, drho_scale)
    if with_walls:
        add_dummy_wall(sc, tank_wall_points(o, size, dx, L, W), 0.5 * dx)
    sc.dt = 0.25 * 1.3 * dx / c0
    return sc


def ensemble_scene(k, config_id=4, n_range=(2000, 60000), nd_max=500, L_range=(0.8, 1.6),
                   alpha_range=(0.01, 0.1), r=0.01, rho_s=2500.0, E=1.0e7, nu_p=0.3, res=0.5,
                   mu=0.4) -> Scene:
    """Scene k of the batched dam-break ensemble.  Draws from sub-stream k of
    stream config_id: fluid count log-uniform in n_range rounded to a cube
    n^3 (dx = 0.4/n, the 0.4 m column), DEM count uniform in {0..nd_max},
    tank length uniform in L_range (0.6 m wide and high), artificial
    viscosity alpha uniform in alpha_range.  The DEM spheres sit on a
    jittered lattice downstream of the column; Hertz-Mindlin contact and
    tank planes as in the coupled dam break."""
    base = 1_000_000 * config_id + 16 * k
    u = uniform(ROOT_SEED, base, 4)
    lo, hi = n_range
    n_f = math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u[0])
    n = max(2, int(round(n_f ** (1.0 / 3.0))))
    n_d = min(nd_max, int(u[1] * (nd_max + 1)))
    L = L_range[0] + (L_range[1] - L_range[0]) * u[2]
    alpha = alpha_range[0] + (alpha_range[1] - alpha_range[0]) * u[3]
    dx = 0.4 / n
    sc = dam_break(dx=dx, L=L, name=f"ensemble_{k}")
    sc.fluid_alpha = np.full(sc.n_fluid, alpha)
    # DEM spheres on a 0.025 m lattice from x = 0.5 downstream, y centred
    sp = 0.025
    nx = max(1, int((L - 0.05 - 0.5) / sp))
    ny = int(0.5 / sp)
    idx = np.arange(n_d)
    pos = np.stack([0.5 + sp * (idx % nx), 0.05 + sp * ((idx // nx) % ny),
                    0.012 + sp * (idx // (nx * ny))], axis=1).astype(float)
    if n_d:
        jit = uniform(ROOT_SEED, base + 4, 3 * n_d, -0.1 * r, 0.1 * r).reshape(n_d, 3)
        add_solid(sc, pos + jit, r, rho_s, material=0)
    Es, Gs = hertz_effective(E, nu_p)
    sc.set_hertz(0, 0, Es, Gs, res, 0.0, 0.0, mu, 0.0, 0.0)
    add_tank_planes(sc, L, 0.6)
    k_c = 0.5 * E * math.pi * r
    dt_dem = 0.2 * math.sqrt(8 * r ** 3 * rho_s / k_c)
    sc.dt = min(sc.dt, dt_dem) if n_d else sc.dt
    sc.alpha = alpha
    return sc


def ensemble(n_scenes=256, **kw) -> list:
    """The batched dam-break ensemble: n_scenes independent scenes (see
    ensemble_scene); run together they step with the smallest dt."""
    return [ensemble_scene(k, **kw) for k in range(n_scenes)]


# --------------------------------------------------------------------------
# the reference's own case driver (driver/cases/damBreak.cpp:14-58)
# --------------------------------------------------------------------------
def dam_break_case(dx=0.025, tank=(8.0, 0.7, 0.8), fluid=(3.5, 0.7, 0.4),
                   clump_origin=(5.2, 0.275, 0.0), cube=0.15, E=0.3e9, cube_density=800.0,
                   floor_point=None, name="dam_break_case") -> Scene:
    """damBreak.cpp as written upstream, with its sizes as parameters (the
    defaults are the reference's): a tank of three wall-dummy layers, a
    water column in the corner (c = 30 m/s, nu = 1e-6), an infinite floor
    plane of material 0, one clump of (cube / dx)^3 touching pebbles of
    material 1 (the 0.15 m cube of 216 pebbles upstream) with the box's
    inertia tensor, Hertz-Mindlin for material pairs (0,1) and (1,1), the
    DEM time step 0.2 sqrt(m / k) and the fluid sub-cycling
    gap = int(dt_f / dt) with dt_f = 0.6 h / (22 sqrt(0.4 g))."""
    L, W, H = tank
    o = np.array([-3 * dx] * 3)
    size = np.array([L + 6 * dx, W + 6 * dx, H + 6 * dx])
    sc = empty_scene(o, size, name=name)
    h = 1.3 * dx
    add_fluid(sc, regular_packed_points((0, 0, 0), fluid, dx), h, 1000.0, 30.0, 1.0e-6)
    r = 0.5 * dx
    add_dummy_wall(sc, tank_wall_points(o, size, dx, L, W), r)
    add_infinite_wall(sc, floor_point if floor_point is not None else (0.5 * L, 0.5 * W, 0.0),
                      (0, 0, 1), 0)
    sc.set_hertz(0, 1, E, 0.3 * E, 0.9, 0, 0, 0.35, 0, 0)
    sc.set_hertz(1, 1, 0.5 * E, 0.15 * E, 0.9, 0, 0, 0.45, 0, 0)
    pts = regular_packed_points(clump_origin, (cube, cube, cube), 2 * r)
    # products in the reference's evaluation order, so the mass and the
    # time step round as they do there
    mass = cube_density * cube * cube * cube
    inertia = (mass / 12.0 * (cube * cube + cube * cube),) * 3 + (0.0, 0.0, 0.0)
    centroid = np.asarray(clump_origin, float) + 0.5 * cube
    add_clump(sc, pts, np.full(len(pts), r), centroid, mass, inertia, material=1)
    k = 0.5 * E * math.pi * r
    mij = 8 * r * r * r * cube_density
    sc.dt = 0.2 * math.sqrt(mij / k)
    dt_f = 0.6 * 1.3 * dx / (22.0 * math.sqrt(0.4 * 9.8))
    sc.gap = max(1, int(dt_f / sc.dt))
    return sc


def dam_break_case_small(name="dam_break_case_small") -> Scene:
    """damBreak.cpp at its own particle spacing, pebble size, materials, time
    step and sub-cycling (gap = 48), in a 1.0 x 0.35 x 0.4 m tank with a
    0.3 m column and the 216-pebble clump 0.035 m from it, so the fluid
    and the clump interact from the first step."""
    return dam_break_case(tank=(1.0, 0.35, 0.4), fluid=(0.3, 0.35, 0.3),
                          clump_origin=(0.31, 0.1, 0.0), name=name)


def immersed_spheres(n_box=(12, 10, 10), dx=0.005, stride=3, r_frac=0.5, rho_s=2500.0,
                     seed_stream=51, name="immersed_spheres") -> Scene:
    """DEM spheres inside a fluid block: every stride-th lattice site of the
    block's interior holds a sphere of radius r_frac dx instead of a fluid
    particle, so each sphere has fluid within 2h on all sides.  Walls of
    dummies around, tank planes for the spheres, Hertz-Mindlin contact."""
    L, W, Hb = (n * dx for n in n_box)
    o = np.array([-3 * dx] * 3)
    size = np.array([L + 6 * dx, W + 6 * dx, 2 * Hb + 6 * dx])
    sc = empty_scene(o, size, name=name)
    pts = regular_packed_points((0, 0, 0), (L, W, Hb), dx)
    ijk = np.rint(pts / dx - 0.5).astype(int)
    inner = np.all((ijk >= 2) & (ijk < np.array(n_box) - 2), axis=1)
    dem = inner & np.all(ijk % stride == 1, axis=1)
    c0 = sound_speed(Hb)
    add_fluid(sc, pts[~dem], 1.3 * dx, RHO0, c0, NU)
    m = sc.n_fluid
    sc.fluid_vel = uniform(ROOT_SEED, seed_stream, 3 * m, -0.05, 0.05).reshape(m, 3)
    sc.fluid_drho = uniform(ROOT_SEED, seed_stream + 1, m, -0.5, 0.5)
    add_dummy_wall(sc, tank_wall_points(o, size, dx, L, W), 0.5 * dx)
    r = r_frac * dx
    add_solid(sc, pts[dem], r, rho_s, material=0)
    Es, Gs = hertz_effective(1.0e7, 0.3)
    sc.set_hertz(0, 0, Es, Gs, 0.5, 0.0, 0.0, 0.4, 0.0, 0.0)
    add_tank_planes(sc, L, W)
    k = 0.5 * 1.0e7 * math.pi * r
    sc.dt = min(0.25 * 1.3 * dx / c0, 0.2 * math.sqrt(8 * r ** 3 * rho_s / k))
    sc.gap = 1
    return sc


def poured(dx=0.005, name="poured"):
    """A scene and the particles added to it after the run started
    (dataBase::addFluid / addCluster / addClump mid-run,
    core/dataBase.cpp:113-117, 226-346).  The scene: spheres inside a fluid
    block in a dummy-walled tank (immersed_spheres), a bonded cluster of
    eight touching spheres above the fluid and two touching spheres sliding
    past each other.  The addition: a fluid block
    above the fluid, a bonded cluster of four touching spheres against the
    first cluster, and a clump of two pebbles.  Returns (scene, addition)."""
    base = immersed_spheres(n_box=(8, 6, 6), dx=dx, name=name)
    r = 0.5 * dx
    d = 2.0 * r * 0.999  # touching neighbours in a cluster
    g = np.array([[i, j, k] for k in range(2) for j in range(2) for i in range(2)], float)
    add_cluster(base, np.array([0.005, 0.01, 0.042]) + d * g, r, 2500.0, material=0)
    # two unbonded touching spheres sliding past each other (a contact with
    # tangential springs)
    n0 = base.n_solid
    add_solid(base, [[0.03, 0.005, 0.042], [0.03 + d, 0.005, 0.042]], r, 2500.0, material=0)
    base.solid_vel[n0:] = [[0.0, 0.01, 0.0], [0.0, -0.01, 0.0]]
    # a bond stiffness of 1e4 N/m: stable at the scene's step
    Eb = 1.0e4 * (2 * r) / (math.pi * r * r)
    base.set_bonded(0, 0, Eb, 1.0, 1.0, 1.0e6, 0.0, 0.0)
    extra = empty_scene(base.domain_origin, base.domain_size, n_material=base.n_material)
    pts = regular_packed_points((0.02, 0.005, 0.045), (3 * dx, 3 * dx, 2 * dx), dx)
    add_fluid(extra, pts, 1.3 * dx, RHO0, sound_speed(6 * dx), NU)
    g4 = np.array([[0, j, k] for k in range(2) for j in range(2)], float)
    add_cluster(extra, np.array([0.005 + 2 * d, 0.01, 0.042]) + d * g4, r, 2500.0, material=0)
    extra.solid_cluster = extra.solid_cluster + int(np.max(base.solid_cluster)) + 1
    peb = np.array([[0.03, 0.02, 0.06], [0.03 + 2 * r, 0.02, 0.06]])
    m = 2 * 4.0 / 3.0 * math.pi * r ** 3 * 2500.0
    I = 0.4 * m * r * r + 0.5 * m * r * r
    add_clump(extra, peb, np.full(2, r), peb.mean(axis=0), m, (I, I, I, 0.0, 0.0, 0.0), material=0)
    return base, extra


def dem_contact_mix(n_side=6, n_clump_side=3, r=0.002, rho_s=2500.0, overlap=0.02, seed_stream=61,
                    name="dem_contact_mix") -> Scene:
    """A DEM scene with every spring of the contact laws loaded from the
    first step: a block of n_side^3 free spheres on a lattice that overlaps
    neighbours by `overlap` of a diameter, with seeded random velocities,
    spins and materials (about 30% material 1, the rest 0), resting on a
    floor plane (material 1); beside it n_clump_side^3 tetrahedral 4-pebble
    clumps (material 1) packed so neighbouring clumps' pebbles overlap by
    the same fraction.  Material pairs: (0,0) Hertz-Mindlin with rolling and
    torsion springs (kr/ks 0.5, kt/ks 0.3, mu_r 0.1, mu_t 0.05), (0,1)
    Hertz-Mindlin, (1,1) the linear spring-dashpot law (LinearContact)."""
    d = 2 * r * (1.0 - overlap)
    span = max(n_side, 2 * n_clump_side) * d * 2
    sc = empty_scene((-0.01, -0.01, -0.01), (span + 0.03, n_side * d + 0.02, span + 0.03), name=name)
    ii, jj, kk = np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij")
    pos = np.stack([ii.ravel() * d + r, jj.ravel() * d + r, kk.ravel() * d + r * (1 - overlap)], axis=1)
    m = len(pos)
    add_solid(sc, pos, r, rho_s, material=0)
    sc.solid_material = (uniform(ROOT_SEED, seed_stream + 2, m) < 0.3).astype(np.int32)
    sc.solid_vel = uniform(ROOT_SEED, seed_stream, 3 * m, -0.05, 0.05).reshape(m, 3)
    sc.solid_angvel = uniform(ROOT_SEED, seed_stream + 1, 3 * m, -50.0, 50.0).reshape(m, 3)
    # clumps: regular tetrahedra of touching pebbles of radius r, on a
    # lattice whose pitch makes neighbouring clumps' pebbles overlap
    edge = 2 * r
    # the nearest pebbles of clumps adjacent along an axis are sqrt((p -
    # sqrt(2) r)^2 + 2 r^2) apart at pitch p: set that to 2 r (1 - overlap)
    pitch = r * (math.sqrt(2.0) + math.sqrt(4.0 * (1.0 - overlap) ** 2 - 2.0))
    x0 = n_side * d + 2 * r
    z0 = r * (math.sqrt(2.0) / 2.0 + 1.0 - overlap)
    m_one = 4.0 / 3.0 * math.pi * r ** 3 * rho_s
    d2 = float(np.sum(tetra_offsets(edge)[0] ** 2))
    Ii = 4 * (0.4 * m_one * r * r) + 4 * m_one * d2 * 2.0 / 3.0
    for c in range(n_clump_side ** 3):
        a, b, e = c % n_clump_side, (c // n_clump_side) % n_clump_side, c // n_clump_side ** 2
        cen = np.array([x0 + pitch * a + r, r + pitch * b, z0 + pitch * e])
        add_clump(sc, cen + tetra_offsets(edge), np.full(4, r), cen, 4 * m_one, (Ii, Ii, Ii, 0, 0, 0),
                  material=1)
    add_infinite_wall(sc, (0, 0, 0), (0, 0, 1), 1)
    Es, Gs = hertz_effective(1.0e7, 0.3)
    sc.set_hertz(0, 0, Es, Gs, 0.5, 0.5, 0.3, 0.4, 0.1, 0.05)
    sc.set_hertz(0, 1, Es, Gs, 0.6, 0.4, 0.2, 0.3, 0.08, 0.04)
    # LinearContact (SPHKernelDEMContactModel.h:214-268): k_n, k_s, k_r, k_t,
    # d_n, d_s, d_r, d_t, mu_s, mu_r, mu_t
    kn = 2.0e3
    sc.set_linear(1, 1, kn, 0.5 * kn, 0.1 * kn, 0.05 * kn, 0.02, 0.01, 0.005, 0.005, 0.35, 0.1, 0.05)
    sc.dt = 1.0e-6
    sc.gap = 1
    return sc


def bonded_bar(n=(10, 3, 3), r=0.002, rho_s=2500.0, pull=0.5, k_bond=1.0e8, sigma_t=1.0e6,
               name="bonded_bar") -> Scene:
    """A bar of touching bonded spheres (one cluster) whose halves move apart
    at +-pull m/s along x, no gravity: the bonds across the middle load in
    tension past sigma_t and break."""
    nx, ny, nz = n
    sc = empty_scene((-0.05, -0.05, -0.05), (0.1 + 2 * nx * r, 0.1, 0.1), name=name)
    sc.gravity = np.zeros(3)
    idx = np.arange(nx * ny * nz)
    ix, iy, iz = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    pos = np.stack([2 * r * ix, 2 * r * iy, 2 * r * iz], axis=1).astype(float)
    vel = np.zeros_like(pos)
    vel[:, 0] = np.where(ix < nx // 2, -pull, pull)
    add_cluster(sc, pos, r, rho_s, material=0)
    sc.solid_vel = vel
    Es, Gs = hertz_effective(1.0e7, 0.3)
    sc.set_hertz(0, 0, Es, Gs, 0.5, 0.0, 0.0, 0.4, 0.0, 0.0)
    Eb = k_bond * (2 * r) / (math.pi * r * r)
    sc.set_bonded(0, 0, Eb, 1.0, 1.0, sigma_t, 0.0, 0.0)
    m = 4.0 / 3.0 * math.pi * r ** 3 * rho_s
    sc.dt = 0.2 * 2.0 / math.sqrt(k_bond / m) / 2.0
    sc.gap = 1
    return sc


def rotating_drum(R=0.03, L=0.03, r=0.002, rho_s=2500.0, omega=10.0, seed_stream=71,
                  with_post=True, name="rotating_drum") -> Scene:
    """DEM spheres in a horizontal drum (a cylinder wall of radius R about
    the x axis, turning at omega) closed by two infinite planes at x = 0
    and x = L; with_post adds a second, fixed cylinder wall of radius R/4
    along y above the drum's axis that spheres thrown up meet from the
    outside.  Spheres on a jittered lattice filling the drum's lower half,
    the outermost pressing on its wall; Hertz-Mindlin with rolling
    friction."""
    sc = empty_scene((-0.01, -R - 0.01, -R - 0.01), (L + 0.02, 2 * R + 0.02, 2 * R + 0.02), name=name)
    d = 2.05 * r
    pts = []
    for x in np.arange(r * 1.2, L - r, d):
        for y in np.arange(-R + 0.98 * r, R - 0.98 * r, d):
            for z in np.arange(-R + 0.98 * r, 0.0, d):
                if y * y + z * z < (R - 0.95 * r) ** 2:   # the outer ones press on the drum
                    pts.append((x, y, z))
    pts = np.asarray(pts, float)
    m = len(pts)
    pts = pts + uniform(ROOT_SEED, seed_stream, 3 * m, -0.01 * r, 0.01 * r).reshape(m, 3)
    add_solid(sc, pts, r, rho_s, material=0)
    Es, Gs = hertz_effective(1.0e7, 0.3)
    sc.set_hertz(0, 0, Es, Gs, 0.5, 0.3, 0.2, 0.4, 0.05, 0.05)
    sc.set_hertz(0, 1, Es, Gs, 0.6, 0.3, 0.2, 0.5, 0.05, 0.05)
    add_infinite_wall(sc, (0, 0, 0), (1, 0, 0), 1)
    add_infinite_wall(sc, (L, 0, 0), (-1, 0, 0), 1)
    add_cylinder_wall(sc, (0, 0, 0), (1, 0, 0), (omega, 0, 0), R, 1)
    if with_post:
        add_cylinder_wall(sc, (0.5 * L, 0, 0.4 * R), (0, 1, 0), (0, 0, 0), 0.25 * R, 1)
    k = 0.5 * 1.0e7 * math.pi * r
    sc.dt = 0.2 * math.sqrt(8 * r ** 3 * rho_s / k)
    sc.gap = 1
    return sc
