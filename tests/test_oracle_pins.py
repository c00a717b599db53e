"""Pins the CPU oracle against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures, and its CUDA sources
cannot be built here (nvcc/thrust absent), so every oracle function is
checked against closed forms, invariants, brute force or an independent
integration — chosen so that a dropped term, a wrong sign or a transposed
index fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
import refmodel
import scenes


# ------------------------------------------------------------- SPH kernel
@pytest.mark.parametrize("h", [0.026, 0.0065, 1.0])
def test_wendland_normalised(h):
    # integral of W over R^3 = 1 (radial quadrature, 4 pi r^2 W dr)
    r = np.linspace(0, 2 * h, 20001)
    w = np.array([oracle.wendland(x, h) for x in r])
    integral = np.trapezoid(4 * np.pi * r * r * w, r)
    assert abs(integral - 1.0) < 1e-7


def test_wendland_peak_and_support():
    h = 0.013
    assert oracle.wendland(0.0, h) == pytest.approx(21.0 / (16.0 * math.pi * h ** 3), rel=1e-15)
    assert oracle.wendland(2 * h, h) == 0.0
    assert oracle.wendland(2.5 * h, h) == 0.0
    assert oracle.wendland(1.999999 * h, h) > 0.0


def test_grad_wendland_matches_finite_difference():
    h = 0.02
    rng = np.random.default_rng(3)
    for _ in range(50):
        rij = rng.uniform(-1.9 * h, 1.9 * h, 3) / math.sqrt(3)
        g = oracle.grad_wendland(rij, h)
        eps = 1e-7 * h
        fd = np.zeros(3)
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            fd[k] = (oracle.wendland(np.linalg.norm(rij + e), h) -
                     oracle.wendland(np.linalg.norm(rij - e), h)) / (2 * eps)
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    assert np.all(oracle.grad_wendland([0, 0, 0], h) == 0)
    assert np.all(oracle.grad_wendland([2.0 * h, 0, 0], h) == 0)


# --------------------------------------------------------- material table
def test_combined_index_closed_form():
    n = 10
    seen = set()
    for a in range(n):
        for b in range(n):
            c = oracle.combined_index(n, a, b)
            assert c == oracle.combined_index(n, b, a)
            i, j = min(a, b), max(a, b)
            # row-major upper triangle enumeration
            expect = sum(n - k for k in range(i)) + (j - i)
            assert c == expect
            seen.add(c)
    assert seen == set(range(n * (n + 1) // 2))
    assert oracle.combined_index(n, n, 0) == -1


# -------------------------------------------------------- grid, hash, sort
def _oracle_for(scene):
    o = oracle.Oracle()
    o.upload(scene)
    o.initialize()
    return o


@pytest.fixture(scope="module")
def cfg0():
    return scenes.dam_break(0.02)


def test_cfg0_sizes_and_grid(cfg0):
    assert cfg0.n_fluid == 8000 and cfg0.n_solid == 32256
    o = _oracle_for(cfg0)
    gs, cs, ncell = o.grid_info()
    assert list(gs) == [33, 13, 13] and ncell == 5578
    origin, size, G, cell, nc = refmodel.grid_params(cfg0)
    assert np.array_equal(cs, cell)


def test_hash_perm_ranges_against_numpy(cfg0):
    o = _oracle_for(cfg0)
    origin, size, G, cell, ncell = refmodel.grid_params(cfg0)
    for species, pos in ((0, cfg0.fluid_pos), (1, cfg0.solid_pos)):
        gr = o.grid(species)
        h, _ = refmodel.point_hash(pos, origin, size, G, cell, ncell)
        assert np.array_equal(gr["hash"], h)
        perm = np.argsort(h, kind="stable")
        assert np.array_equal(gr["perm"], perm)
        st, en = refmodel.cell_ranges(h[perm], ncell)
        assert np.array_equal(gr["start"], st) and np.array_equal(gr["end"], en)


def test_out_of_domain_goes_to_spare_cell():
    sc = scenes.random_fluid_box(n=0, box=(0.06, 0.05, 0.04), dx=0.005, with_walls=False)
    sc.fluid_pos = sc.fluid_pos.copy()
    sc.fluid_pos[0] = [10.0, 0.0, 0.0]
    sc.fluid_pos[1] = sc.domain_origin + sc.domain_size  # max bound is exclusive
    o = _oracle_for(sc)
    gr = o.grid(0)
    ncell = o.grid_info()[2]
    assert gr["hash"][0] == ncell - 1 and gr["hash"][1] == ncell - 1
    assert gr["perm"][-2:].tolist() == [0, 1]  # stable: ascending index in the spare cell


# ------------------------------------------------------------------ lists
def _small_scenes():
    a = scenes.random_fluid_box(box=(0.05, 0.04, 0.04), dx=0.005, seed_stream=11, h_jitter=0.15)
    b = scenes.dam_break(0.04)
    return [a, b]


@pytest.mark.parametrize("k", [0, 1])
def test_fluid_lists_against_brute_force(k):
    sc = _small_scenes()[k]
    o = _oracle_for(sc)
    ffi, ffj, fsi, fsa = refmodel.brute_fluid_lists(sc)
    a, b, cnt = o.list(0)
    assert np.array_equal(a, ffi) and np.array_equal(b, ffj)
    assert np.array_equal(np.bincount(ffi, minlength=sc.n_fluid), cnt)
    a, b, _ = o.list(1)
    assert np.array_equal(a, fsi) and np.array_equal(b, fsa)


def test_transpose_is_ascending_pointed():
    sc = _small_scenes()[0]
    o = _oracle_for(sc)
    a, b, _ = o.list(0)
    idx, rs, re = o.transpose(0)
    # stable sort of pair ids by pointing particle
    assert np.array_equal(idx, np.argsort(b, kind="stable"))
    for j in range(sc.n_fluid):
        seg = idx[rs[j]:re[j]] if rs[j] >= 0 else []
        assert np.all(b[seg] == j)
        assert np.all(np.diff(a[seg]) > 0)


def test_solid_contact_list_against_brute_force():
    sc = scenes.granular_column(n_sphere=300, n_clump=30, n_bonded=125)
    o = _oracle_for(sc)
    I, J = refmodel.brute_solid_list(sc)
    a, b, cnt = o.list(2)
    assert np.array_equal(a, I) and np.array_equal(b, J)
    assert len(a) > 0


# -------------------------------------------------------------- rotations
def _quat_matrix(q):
    q0, q1, q2, q3 = q
    return np.array([
        [1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
        [2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1)],
        [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)]])


def test_rotate_inverse_inertia_is_R_I_RT():
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        A = rng.normal(size=(3, 3))
        I = A @ A.T + 3 * np.eye(3)
        invI = np.linalg.inv(I)
        v6 = [invI[0, 0], invI[1, 1], invI[2, 2], invI[0, 1], invI[0, 2], invI[1, 2]]
        out = oracle.rotate_inverse_inertia(q, v6)
        R = _quat_matrix(q)
        M = R @ invI @ R.T
        expect = [M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]]
        assert np.allclose(out, expect, rtol=1e-12, atol=1e-14)
    ident = oracle.rotate_inverse_inertia([1, 0, 0, 0], [1, 2, 3, 0.1, 0.2, 0.3])
    assert np.allclose(ident, [1, 2, 3, 0.1, 0.2, 0.3], rtol=0, atol=0)


def test_quaternion_rotate_unit_and_first_order():
    rng = np.random.default_rng(7)
    q = np.array([1.0, 0, 0, 0])
    w = np.array([0.3, -0.2, 0.5])
    dt = 1e-4
    for _ in range(2000):
        q = oracle.quaternion_rotate(q, w, dt)
        assert abs(np.linalg.norm(q) - 1) < 1e-14
    # total rotation angle |w| T about w: compare with the exact rotation
    T = 2000 * dt
    th = np.linalg.norm(w) * T
    k = w / np.linalg.norm(w)
    exact = np.r_[math.cos(th / 2), math.sin(th / 2) * k]
    assert np.allclose(q, exact, atol=1e-8)
    # world-frame rate: R(q_t) = exp([w]x t) R(q_0)
    R = _quat_matrix(q)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    Rexp = np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K
    assert np.allclose(R, Rexp, atol=1e-7)


def test_rodrigues_rotation():
    rng = np.random.default_rng(9)
    for _ in range(30):
        v = rng.normal(size=3)
        ang = rng.normal(size=3)
        out = oracle.rotate_vector(v, ang)
        th = np.linalg.norm(ang)
        k = ang / th
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K
        assert np.allclose(out, R @ v, rtol=1e-12, atol=1e-13)
        assert abs(np.linalg.norm(out) - np.linalg.norm(v)) < 1e-12
    v = np.array([1.0, 2, 3])
    assert np.array_equal(oracle.rotate_vector(v, [1e-11, 0, 0]), v)


# ------------------------------------------------------------ contact laws
HERTZ = [5.4945e6, 1.131e6, 0.0, 0.0, 0.0, 0.4, 0.0, 0.0]  # E*, G*, D, kr/ks, kt/ks, mu...


def test_hertz_normal_force_law_and_exponent():
    # F_n = 4/3 E* sqrt(R) d^{3/2}: doubling the overlap multiplies F by 2^{3/2}
    n = np.array([0, 0, 1.0])
    R = 0.005
    F1, _, _ = oracle.contact(0, [0, 0, 0], [0, 0, 0], n, 1e-5, 0.01, R, 1e-5, HERTZ)
    F2, _, _ = oracle.contact(0, [0, 0, 0], [0, 0, 0], n, 2e-5, 0.01, R, 1e-5, HERTZ)
    assert F1[2] > 0 and abs(F2[2] / F1[2] - 2 ** 1.5) < 1e-12
    assert F1[2] == pytest.approx(4 / 3 * HERTZ[0] * math.sqrt(R) * 1e-5 ** 1.5, rel=1e-12)


def test_coulomb_cap_and_spring_projection():
    rng = np.random.default_rng(11)
    p = list(HERTZ)
    p[2] = 0.2  # damping
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v = rng.normal(size=3) * 0.5
        sp = rng.normal(size=(3, 3)) * 1e-5
        F, T, sp2 = oracle.contact(0, v, [0, 0, 0], n, 1e-5, 0.01, 0.005, 1e-5, p, sp)
        Fn = np.dot(F, n) * n
        Ft = F - Fn
        vn = np.dot(v, n) * n
        # the normal force includes damping; the cap uses |F_n| of that force
        assert np.linalg.norm(Ft) <= 0.4 * np.linalg.norm(Fn) * (1 + 1e-9) + 1e-12
    # projection onto the tangent plane keeps the spring magnitude (mu large,
    # zero velocity, no damping)
    p = list(HERTZ)
    p[5] = 1e9
    n = np.array([0, 0, 1.0])
    s0 = np.array([[1e-6, 0, 1e-6], [0, 0, 0], [0, 0, 0]])
    _, _, s1 = oracle.contact(0, [0, 0, 0], [0, 0, 0], n, 1e-5, 0.01, 0.005, 1e-5, p, s0)
    assert abs(s1[0, 2]) < 1e-20
    assert np.linalg.norm(s1[0]) == pytest.approx(np.linalg.norm(s0[0]), rel=1e-12)


def test_contact_releases_springs_when_separated():
    sp = np.ones((3, 3))
    F, T, s = oracle.contact(0, [1, 0, 0], [0, 0, 1], [0, 0, 1.0], -1e-6, 0.01, 0.005, 1e-5,
                             HERTZ, sp)
    assert np.all(F == 0) and np.all(T == 0) and np.all(s == 0)


def _head_on_rebound(model_params, v0, r, rho, dt, model=0):
    """Two equal spheres approaching head-on, integrated by the oracle's
    contact law with symplectic Euler; returns the rebound speed ratio."""
    m = 4 / 3 * math.pi * r ** 3 * rho
    meff = m / 2
    x = np.array([-r, r])  # touching at t=0
    v = np.array([v0, -v0])
    n = np.array([1.0, 0, 0])
    for _ in range(200000):
        d = 2 * r - (x[1] - x[0])
        vrel = v[0] - v[1]  # i = 0 (left) approaches j = 1 while vrel > 0
        if d <= 0 and vrel < 0:
            break
        # relative velocity at the contact v_i - v_j, normal n_ij = unit(r_i - r_j) = -x
        F, _, _ = oracle.contact(model, [vrel, 0, 0], [0, 0, 0], -n, d, meff, r / 2, dt,
                                 model_params)
        a0 = F[0] / m
        v[0] += a0 * dt
        v[1] -= a0 * dt
        x += v * dt
    return (v[1] - v[0]) / (2 * v0)


def test_hertz_restitution_matches_ode():
    # independent integration of the same damped Hertz law with scipy
    from scipy.integrate import solve_ivp
    E, G, res = 5.4945e6, 1.131e6, 0.5
    lr = math.log(res)
    D = -lr / math.sqrt(lr * lr + math.pi ** 2)
    r, rho, v0 = 0.01, 2500.0, 0.5
    m = 4 / 3 * math.pi * r ** 3 * rho
    meff, R = m / 2, r / 2
    params = [E, G, D, 0, 0, 0, 0, 0]

    def rhs(t, y):
        d, dd = y
        kn = 4 / 3 * E * math.sqrt(R * max(d, 0))
        c = 2 * math.sqrt(5 / 6) * D * math.sqrt(meff * kn)
        F = kn * d + c * dd if d > 0 else 0.0
        return [dd, -F / meff]

    ev = lambda t, y: y[0] + 1e-15 if t > 0 else 1.0
    ev.terminal, ev.direction = True, -1
    sol = solve_ivp(rhs, [0, 1.0], [0.0, 2 * v0], events=ev, rtol=1e-11, atol=1e-16,
                    max_step=1e-6)
    e_ode = -sol.y_events[0][0][1] / (2 * v0)
    e_orc = _head_on_rebound(params, v0, r, rho, 2e-8)
    assert abs(e_orc - e_ode) < 2e-3
    assert 0.3 < e_orc < 0.7


# ----------------------------------------------------------------- bonds
BOND = [1.0, 1.0e10, 1.0, 1.0e6, 0.0, 0.0]  # gamma, E, kn/ks, sigma_t, C, mu


def test_bond_small_strain_linear_response():
    ra = rb = 0.002
    n = np.array([0, 0, 1.0])
    dt = 1e-7
    vn = -1e-3  # approaching -> compression increment
    st, on = oracle.parallel_bond(np.zeros(8), n, n, [0, 0, vn], [0, 0, 0], [0, 0, 0], ra, rb,
                                  dt, BOND)
    A = math.pi * ra * ra
    kA = BOND[1] / (ra + rb)
    assert st[0] == pytest.approx(-(vn * dt) * kA * A, rel=1e-12)
    assert on == 1


def test_bond_breaks_at_tensile_strength():
    ra = rb = 0.002
    A = math.pi * ra * ra
    n = np.array([0, 0, 1.0])
    # tension Fn such that sigma = -Fn/A crosses sigma_t
    st = np.zeros(8)
    st[0] = -(BOND[3] * A) * 0.999
    _, on = oracle.parallel_bond(st, n, n, [0, 0, 0], [0, 0, 0], [0, 0, 0], ra, rb, 1e-7, BOND)
    assert on == 1
    st[0] = -(BOND[3] * A) * 1.001
    _, on = oracle.parallel_bond(st, n, n, [0, 0, 0], [0, 0, 0], [0, 0, 0], ra, rb, 1e-7, BOND)
    assert on == 0


def test_bond_rotation_preserves_shear_norm():
    ra = rb = 0.002
    st = np.zeros(8)
    st[2:5] = [3.0, -1.0, 0.0]
    st[5:8] = [0.0, 2.0, 0.0]
    n0 = np.array([0, 0, 1.0])
    n1 = np.array([math.sin(0.01), 0, math.cos(0.01)])
    p = list(BOND)
    p[3] = 0
    out, on = oracle.parallel_bond(st, n0, n1, [0, 0, 0], [0, 0, 0.3], [0, 0, 0.3], ra, rb, 1e-3, p)
    assert np.linalg.norm(out[2:5]) == pytest.approx(np.linalg.norm(st[2:5]), rel=1e-12)
    assert np.linalg.norm(out[5:8]) == pytest.approx(np.linalg.norm(st[5:8]), rel=1e-12)


# ---------------------------------------------------- fluid invariants
def _periodic_free_lattice(n=12, dx=0.01):
    """A cube of fluid at rest, no walls, no gravity."""
    sc = scenes.empty_scene((-0.05, -0.05, -0.05), (0.25, 0.25, 0.25))
    sc.gravity = np.zeros(3)
    pts = scenes.regular_packed_points((0, 0, 0), (n * dx,) * 3, dx)
    scenes.add_fluid(sc, pts, 1.3 * dx, 1000.0, 20.0, 1e-6)
    return sc, n, dx


def test_lattice_at_rest_interior_is_static():
    sc, n, dx = _periodic_free_lattice()
    o = _oracle_for(sc)
    o.step(1, 1e-5)
    f = o.fluid()
    p = sc.fluid_pos
    lo, hi = 3 * dx, (n - 3) * dx  # > 2h = 2.6 dx from the free surface
    inner = np.all((p > lo) & (p < hi), axis=1)
    assert inner.sum() > 50
    assert np.abs(f["drho"][inner]).max() < 1e-12
    assert np.abs(f["acc"][inner]).max() < 1e-8


def test_density_reinit_uniform_lattice_ratio_one():
    sc, n, dx = _periodic_free_lattice()
    o = _oracle_for(sc)
    o.fluid_integrate(1e-5, 1, 1)   # stores sumW0
    o.fluid_integrate(1e-5, 5, 1)   # re-init with the same geometry
    f = o.fluid()
    # rho0 * sumW/sumW0 = rho0 exactly; max(rho*, rho0) - rho0 >= dRho*
    assert np.all(f["drho"] >= -1e-12)


def test_fluid_momentum_balance_without_walls():
    # pair terms are antisymmetric in m_i a_i: total momentum rate = M g
    sc = scenes.random_fluid_box(box=(0.05, 0.04, 0.04), dx=0.005, seed_stream=21,
                                 with_walls=False, vel_scale=0.2, drho_scale=2.0)
    o = _oracle_for(sc)
    o.step(2, 1e-6)
    f = o.fluid()
    m = (np.asarray(sc.fluid_h) / 1.3) ** 3 * np.asarray(sc.fluid_rho0)
    P = (m[:, None] * f["acc"]).sum(axis=0)
    M = m.sum()
    assert np.allclose(P, M * sc.gravity, rtol=0, atol=1e-9 * M * 9.81)


def test_two_particle_symmetric_compression():
    # equal particles approaching head-on: the Riemann interface velocity is
    # 0, so d rho_i/dt = 2 rho V_j u dW/dx  (> 0 for approach)
    sc = scenes.empty_scene((-0.1, -0.1, -0.1), (0.2, 0.2, 0.2))
    sc.gravity = np.zeros(3)
    dx = 0.01
    scenes.add_fluid(sc, [[-0.5 * dx, 0, 0], [0.5 * dx, 0, 0]], 1.3 * dx, 1000.0, 20.0, 1e-6)
    u = 0.1
    sc.fluid_vel = np.array([[u, 0, 0], [-u, 0, 0]])
    o = _oracle_for(sc)
    dt = 1e-7
    o.fluid_integrate(dt, 2, 1)
    f = o.fluid()
    h = 1.3 * dx
    m = (h / 1.3) ** 3 * 1000.0
    V = m / 1000.0
    g = oracle.grad_wendland([-dx, 0, 0], h)  # r_0 - r_1
    drdt = 2 * 1000.0 * V * np.dot([u, 0, 0], g)
    # two half steps of dt/2 with v* = 0 (velocities change by O(dt))
    assert f["drho"][0] == pytest.approx(drdt * dt, rel=1e-4)
    assert drdt > 0
    assert f["drho"][1] == pytest.approx(f["drho"][0], rel=1e-12)


# ------------------------------------------ the Riemann interface states
# The pair terms of continuity and momentum (integrate.cu:100-101, 211-212)
# are the Riemann-based WCSPH of Zhang, Hu & Adams (J. Comput. Phys. 340,
# 2017): along the axis s = -e_ij from particle i (left state) to j (right
# state) each pair solves the linearised (acoustic) Riemann problem with
# impedance Z = rho c (c the pair's mean sound speed).  Its star state
# follows from the two characteristic relations across the acoustic waves,
#     p* + Z_L u* = p_L + Z_L u_L      (the left-going wave)
#     p* - Z_R u* = p_R - Z_R u_R      (the right-going wave),
# solved below as a 2x2 linear system (not a closed form).  Then
#   continuity  drho_i/dt = 2 rho_i V_j (v_i - v*) . grad W_ij with
#               v* = mean (rho-weighted) velocity with its s-component
#               replaced by u* (the wave speed is the full c here);
#   momentum    dv_i/dt = -2 (V_j / rho_i) p* grad W_ij (+ viscous), with
#               p* solved at the low-dissipation wave speed
#               beta = min(3 max(u_L - u_R, 0), c): the full acoustic
#               p* in strong compression, the impedance-weighted pressure
#               with no velocity term when the pair separates.
def _acoustic_star(uL, uR, pL, pR, rhoL, rhoR, w):
    """Star state (u*, p*) of the acoustic Riemann problem with impedances
    Z = rho w.  The unknowns are p* and the momentum-flux jump w u*, so the
    system stays regular as the wave speed w goes to 0 (u* is then
    undetermined and returned as nan)."""
    A = np.array([[1.0, rhoL], [1.0, -rhoR]])
    b = np.array([pL + rhoL * w * uL, pR - rhoR * w * uR])
    p, wu = np.linalg.solve(A, b)
    return (wu / w if w > 0 else float("nan")), p


def _riemann_pair(ri, rj, vi, vj, hi, hj, rho0i, rho0j, dri, drj, c, nu):
    """A two-particle scene (no gravity, no walls) and the rates the
    Riemann WCSPH assigns to particle 0 from the acoustic star state."""
    sc = scenes.empty_scene((-0.2, -0.2, -0.2), (0.4, 0.4, 0.4))
    sc.gravity = np.zeros(3)
    scenes.add_fluid(sc, [ri], hi, rho0i, c, nu)
    scenes.add_fluid(sc, [rj], hj, rho0j, c, nu)
    sc.fluid_vel = np.array([vi, vj], float)
    sc.fluid_drho = np.array([dri, drj], float)
    ri, rj, vi, vj = (np.asarray(x, float) for x in (ri, rj, vi, vj))
    rhoi, rhoj = rho0i + dri, rho0j + drj
    Pi, Pj = c * c * dri, c * c * drj
    Vj = (hj / 1.3) ** 3 * rho0j / rhoj
    rij = ri - rj
    s = -rij / np.linalg.norm(rij)
    gW = oracle.grad_wendland(rij, 0.5 * (hi + hj))
    uL, uR = float(vi @ s), float(vj @ s)
    # continuity: the full acoustic wave speed
    ustar, _ = _acoustic_star(uL, uR, Pi, Pj, rhoi, rhoj, c)
    vbar = (rhoi * vi + rhoj * vj) / (rhoi + rhoj)
    vstar = vbar - (vbar @ s) * s + ustar * s
    drdt = 2 * rhoi * Vj * float((vi - vstar) @ gW)
    # momentum: the low-dissipation wave speed
    beta = min(3.0 * max(uL - uR, 0.0), c)
    _, pstar = _acoustic_star(uL, uR, Pi, Pj, rhoi, rhoj, beta)
    acc = -2.0 * Vj / rhoi * pstar * gW
    return sc, drdt, acc, beta


def _rates(sc, dt=1e-13):
    """d rho/dt and the acceleration of every particle from one fluid step of
    length dt: both continuity half steps and the momentum pass see the
    initial state to O(dt)."""
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize(compute_normals=False)
    o.fluid_integrate(dt, 2, 1)      # no density re-init on step 2
    f = o.fluid()
    o.close()
    return (f["drho"] - np.asarray(sc.fluid_drho)) / dt, f["acc"]


# Unequal densities (different rho0 and dRho), pressures, smoothing lengths
# and approach speeds, with velocity components off the pair axis: strong
# compression (beta = c), weak compression (0 < beta < c) and separation
# (beta = 0).  nu is tiny: the viscous term (pinned separately) is
# negligible here.
_RIEMANN_CASES = {
    # name: (r_j - r_i direction, v_i, v_j, h_i/dx, h_j/dx, rho0_i, rho0_j, dRho_i, dRho_j)
    "strong_compression": ((1.0, 0.3, -0.2), (9.0, 1.0, -2.0), (-4.0, 0.5, 1.5),
                           1.3, 1.2, 1000.0, 820.0, 3.5, -1.25),
    "weak_compression": ((0.2, -1.0, 0.4), (0.3, -1.1, 0.7), (0.1, 0.9, -0.2),
                         1.25, 1.35, 950.0, 1100.0, -2.0, 4.0),
    "separation": ((-0.5, 0.1, 1.0), (0.4, -0.2, -0.9), (-0.3, 0.6, 1.1),
                   1.3, 1.3, 1000.0, 1000.0, 1.75, -0.5),
}


@pytest.mark.parametrize("case", list(_RIEMANN_CASES))
def test_riemann_pair_terms_match_the_acoustic_star_state(case):
    d, vi, vj, hi, hj, rho0i, rho0j, dri, drj = _RIEMANN_CASES[case]
    dx, c = 0.01, 20.0
    ri = np.zeros(3)
    rj = 1.1 * dx * _unit(d)
    sc, drdt, acc, beta = _riemann_pair(ri, rj, vi, vj, hi * dx, hj * dx, rho0i, rho0j,
                                        dri, drj, c, 1e-14)
    {"strong_compression": lambda: beta == c,
     "weak_compression": lambda: 0 < beta < c,
     "separation": lambda: beta == 0.0}[case]() or pytest.fail(f"regime of {case}: beta={beta}")
    rate, a = _rates(sc)
    assert abs(drdt) > 1.0 and np.linalg.norm(acc) > 1.0
    assert rate[0] == pytest.approx(drdt, rel=1e-6)
    assert np.abs(a[0] - acc).max() < 1e-7 * np.linalg.norm(acc)


def test_riemann_terms_swap_left_and_right():
    # particle 1 sees the same pair with left and right exchanged: its
    # rates are the acoustic star state of the mirrored problem
    d, vi, vj, hi, hj, rho0i, rho0j, dri, drj = _RIEMANN_CASES["strong_compression"]
    dx, c = 0.01, 20.0
    rj = 1.1 * dx * _unit(d)
    sc, _, _, _ = _riemann_pair(np.zeros(3), rj, vi, vj, hi * dx, hj * dx, rho0i, rho0j,
                                dri, drj, c, 1e-14)
    _, drdt1, acc1, _ = _riemann_pair(rj, np.zeros(3), vj, vi, hj * dx, hi * dx, rho0j, rho0i,
                                      drj, dri, c, 1e-14)
    rate, a = _rates(sc)
    assert rate[1] == pytest.approx(drdt1, rel=1e-6)
    assert np.abs(a[1] - acc1).max() < 1e-7 * np.linalg.norm(acc1)


def test_viscous_term_is_the_laplacian_of_a_quadratic_shear():
    # v_x = k z^2 in a lattice at rest (no pressure, no gravity): the viscous
    # term approximates nu lap v = 2 nu k in the interior.  The velocities
    # are small so the Riemann dissipation (quadratic in them) is negligible.
    sc, n, dx = _periodic_free_lattice(n=16, dx=0.01)
    nu, k = 1e-2, 1e-4
    p = np.asarray(sc.fluid_pos)
    z0 = 0.5 * n * dx
    vel = np.zeros_like(p)
    vel[:, 0] = k * (p[:, 2] - z0) ** 2
    sc.fluid_vel = vel
    sc.fluid_nu = np.full(sc.n_fluid, nu)
    o = _oracle_for(sc)
    o.fluid_integrate(1e-9, 2, 1)
    a = o.fluid()["acc"]
    lo, hi = 3 * dx, (n - 3) * dx
    inner = np.all((p > lo) & (p < hi), axis=1)
    assert inner.sum() > 500
    ax = a[inner, 0]
    assert np.mean(ax) == pytest.approx(2 * nu * k, rel=0.05)
    assert np.std(ax) < 0.05 * 2 * nu * k
    assert np.abs(a[inner, 1:]).max() < 1e-3 * 2 * nu * k


def test_hydrostatic_column_settles_to_rho_g_depth():
    # a water column at rest in a walled tank, released with no pressure:
    # once the release has spread out, the time-averaged pressure c^2 dRho
    # of each interior layer is rho0 g (depth below the free surface), to
    # the SPH truncation error of the free surface (a few %)
    dx, Hc = 0.01, 0.1
    sc = scenes.dam_break(dx=dx, L=0.1, W=0.1, H=0.15, column=(0.1, 0.1, Hc), name="hydrostatic")
    sc.fluid_nu = np.full(sc.n_fluid, 2e-3)
    o = _oracle_for(sc)
    c = float(sc.fluid_c[0])
    o.run(600, sc.dt)
    layers = (0.02, 0.04, 0.06, 0.08)
    ratio = {lo: [] for lo in layers}
    step = 601
    for _ in range(40):               # 800 more steps, about 26 acoustic crossings
        o.run(20, sc.dt, first=step)
        step += 20
        f = o.fluid()
        P = c * c * f["drho"]
        z = f["pos"][:, 2]
        depth = np.max(z) + 0.5 * dx - z
        xy = f["pos"][:, :2]
        core = np.all((xy > 2 * dx) & (xy < 0.1 - 2 * dx), axis=1)
        for lo in layers:
            sel = core & (depth > lo) & (depth < lo + dx)
            ratio[lo].append(np.mean(P[sel]) / (1000.0 * 9.81 * np.mean(depth[sel])))
    o.close()
    mean = {lo: float(np.mean(v)) for lo, v in ratio.items()}
    for lo in layers:
        assert mean[lo] == pytest.approx(1.0, abs=0.12), mean


# ----------------------------------------------------- DEM integration
def test_sphere_resting_on_plane_static_overlap():
    E, G = 5.4945e6, 1.131e6
    r, rho = 0.01, 2500.0
    sc = scenes.empty_scene((-0.1, -0.1, -0.05), (0.2, 0.2, 0.2))
    scenes.add_solid(sc, [[0, 0, r * 0.999]], r, rho)
    sc.set_hertz(0, 0, E, G, 0.05, 0, 0, 0.4, 0, 0)
    scenes.add_infinite_wall(sc, (0, 0, 0), (0, 0, 1), 0)
    o = _oracle_for(sc)
    dt = 2e-5
    for s in range(1, 6001):
        o.step(s, dt)
    z = o.solid()["pos"][0, 2]
    m = 4 / 3 * math.pi * r ** 3 * rho
    d_expect = (m * 9.81 / (4 / 3 * E * math.sqrt(r))) ** (2 / 3)
    assert r - z == pytest.approx(d_expect, rel=2e-3)


def test_torque_free_clump_keeps_angular_velocity():
    sc = scenes.empty_scene((-0.1, -0.1, -0.1), (0.2, 0.2, 0.2))
    sc.gravity = np.zeros(3)
    r = 0.002
    off = scenes.tetra_offsets(2 * r)
    m1 = 4 / 3 * math.pi * r ** 3 * 2500
    I = 4 * 0.4 * m1 * r * r + 4 * m1 * float(np.sum(off[0] ** 2)) * 2 / 3
    w = np.array([1.0, -2.0, 0.5])
    scenes.add_clump(sc, off, np.full(4, r), np.zeros(3), 4 * m1, (I, I, I, 0, 0, 0), angvel=w)
    o = _oracle_for(sc)
    for s in range(1, 51):
        o.step(s, 1e-5)
    c = o.clump()
    assert np.allclose(c["angvel"][0], w, rtol=0, atol=1e-12)
    assert np.allclose(c["centroid"][0], 0, atol=1e-15)


# ------------------------------------------------------ artificial viscosity
def _av_accel(sc, alpha):
    """The artificial-viscosity part of the momentum rate: acc with the term
    minus acc without it, from the same state (it enters only the momentum
    pass, after the first half-step of continuity and drift)."""
    out = []
    for flags in (0, dsph_flags()["ART_VISCOSITY"]):
        sc.fluid_alpha = np.full(sc.n_fluid, alpha)
        o = oracle.Oracle()
        o.set_flags(flags)
        o.upload(sc)
        o.initialize()
        o.step(2, 1e-6)
        out.append(o.fluid()["acc"])
        o.close()
    return out[1] - out[0]


def dsph_flags():
    import dsph
    return {"ART_VISCOSITY": dsph.FLAG_ART_VISCOSITY}


def test_artificial_viscosity_conserves_momentum_and_dissipates():
    sc = scenes.random_fluid_box(box=(0.05, 0.04, 0.04), dx=0.005, seed_stream=23,
                                 with_walls=False, vel_scale=0.2, drho_scale=2.0)
    sc.gravity = np.zeros(3)
    a = _av_accel(sc, 0.05)
    m = (np.asarray(sc.fluid_h) / 1.3) ** 3 * np.asarray(sc.fluid_rho0)
    scale = np.abs(m[:, None] * a).sum()
    assert scale > 0
    # pair terms are equal and opposite (uniform alpha and mass)
    assert np.abs((m[:, None] * a).sum(axis=0)).max() < 1e-12 * scale
    # the work done on the particles is negative: sum_pairs -|F|(v.r)^2 <= 0
    v = np.asarray(sc.fluid_vel)
    assert (m[:, None] * a * v).sum() < 0


def test_artificial_viscosity_is_linear_in_alpha():
    sc = scenes.random_fluid_box(box=(0.04, 0.03, 0.03), dx=0.005, seed_stream=24,
                                 with_walls=False, vel_scale=0.2, drho_scale=2.0)
    a1, a2 = _av_accel(sc, 0.03), _av_accel(sc, 0.06)
    assert np.abs(a2 - 2 * a1).max() < 1e-9 * np.abs(a2).max()


def test_artificial_viscosity_slows_an_approaching_pair():
    sc = scenes.empty_scene((-0.1, -0.1, -0.1), (0.2, 0.2, 0.2))
    sc.gravity = np.zeros(3)
    dx = 0.01
    scenes.add_fluid(sc, [[-0.5 * dx, 0, 0], [0.5 * dx, 0, 0]], 1.3 * dx, 1000.0, 20.0, 1e-6)
    sc.fluid_vel = np.array([[0.1, 0, 0], [-0.1, 0, 0]])
    a = _av_accel(sc, 0.1)
    assert a[0, 0] < 0 < a[1, 0]          # each pushed back along its approach
    assert np.abs(a[:, 1:]).max() < 1e-12 * abs(a[0, 0])


# ------------------------------------------- fluid–wall terms and normals
# The fluid–solid terms of continuity and momentum (integrate.cu:139-158,
# 251-277) are a Riemann solve against a mirrored state: a wall particle a
# of radius r_a acts on fluid particle i as a fluid particle at r_a would,
# with h = 2.6 r_a, velocity 2 v_a - v_i, pressure P_a, density P_a/c^2 +
# rho0_i and viscosity gamma_i, except that its volume is (2 r_a)^3 instead
# of m/rho.  With the fluid particle on the wall normal through a
# (e_ia = n_a) the two terms differ only by that volume ratio, so the
# fluid–fluid term (pinned above) pins the wall term.
_WALL_DX = 0.01
_H = 1.3 * _WALL_DX
_RHO0, _C = 1000.0, 20.0


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def _wall_scene(n, d, v_i, v_a, drho_i, nu, g):
    """Fluid particle i at distance d from wall particle a along the normal n."""
    sc = scenes.empty_scene((0, 0, 0), (0.1, 0.1, 0.1))
    sc.gravity = np.asarray(g, float)
    p_i = np.array([0.05, 0.05, 0.05])
    scenes.add_fluid(sc, [p_i], _H, _RHO0, _C, nu)
    sc.fluid_vel = np.array([v_i], float)
    sc.fluid_drho = np.array([drho_i])
    scenes.add_dummy_wall(sc, [p_i - d * n], _H / 2.6)
    sc.solid_vel = np.array([v_a], float)
    sc.solid_normal = np.array([n], float)
    return sc, p_i


def _mirror_scene(n, d, v_i, v_a, drho_i, drho_j, nu, g):
    """The same i with a fluid particle j in place of the wall particle."""
    sc = scenes.empty_scene((0, 0, 0), (0.1, 0.1, 0.1))
    sc.gravity = np.asarray(g, float)
    p_i = np.array([0.05, 0.05, 0.05])
    rho_i, rho_j = _RHO0 + drho_i, _RHO0 + drho_j
    scenes.add_fluid(sc, [p_i], _H, _RHO0, _C, nu)
    # gamma_j = nu_j rho_j equal to gamma_i
    scenes.add_fluid(sc, [p_i - d * n], _H, _RHO0, _C, nu * rho_i / rho_j)
    sc.fluid_vel = np.array([v_i, 2 * np.asarray(v_a) - np.asarray(v_i)], float)
    sc.fluid_drho = np.array([drho_i, drho_j])
    return sc


def _fluid_step(sc, dt):
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize(compute_normals=False)
    o.fluid_integrate(dt, 2, 1)      # no density re-init on step 2
    f = o.fluid()
    fs = o.fs_force() if sc.n_solid else None
    o.close()
    return f, fs


_MIRROR_CASES = [
    # normal, distance, v_i, v_a, dRho_i, nu, gravity
    ((0.3, -0.2, 0.93), 0.8, (0.05, -0.1, 0.2), (0.01, 0.02, -0.03), 1.5, 1e-3, (0, 0, 0)),
    ((0, 0, 1), 1.3, (0.0, 0.0, -0.3), (0.0, 0.0, 0.0), -0.7, 1e-6, (0, 0, 0)),
    # gravity along n: the wall pressure takes its hydrostatic increment
    ((0, 0, 1), 0.6, (0.1, 0.0, -0.2), (0.0, 0.05, 0.0), 0.4, 1e-3, (0, 0, 9.81)),
]


@pytest.mark.parametrize("case", _MIRROR_CASES)
def test_wall_term_is_a_mirrored_fluid_particle(case):
    n, d, v_i, v_a, drho_i, nu, g = case
    n = _unit(n)
    d *= _WALL_DX
    rho_i = _RHO0 + drho_i
    # P_a = P_i + rho_i max(0, g.n) (r_ia.n) (integrate.cu:152), as a density
    drho_j = drho_i + rho_i * max(0.0, float(np.dot(g, n))) * d / _C ** 2
    dt = 1e-10
    fa, _ = _fluid_step(_wall_scene(n, d, v_i, v_a, drho_i, nu, g)[0], dt)
    fb, _ = _fluid_step(_mirror_scene(n, d, v_i, v_a, drho_i, drho_j, nu, g), dt)
    vol = (_RHO0 + drho_j) / _RHO0    # (2 r_a)^3 over the mirror particle's m/rho
    rate_a = (fa["drho"][0] - drho_i) / dt
    rate_b = (fb["drho"][0] - drho_i) / dt
    assert abs(rate_b) > 1.0
    assert rate_a == pytest.approx(vol * rate_b, rel=1e-6)
    acc_a = fa["acc"][0] - g
    acc_b = fb["acc"][0] - g
    assert np.linalg.norm(acc_b) > 1.0
    assert np.abs(acc_a - vol * acc_b).max() < 1e-7 * np.linalg.norm(acc_b)


def test_wall_term_is_galilean_invariant():
    # the mirror velocity 2 v_a - v_i moves with the frame
    n = _unit((0.2, 0.1, 1.0))
    u = np.array([0.3, -0.4, 0.25])
    v_i, v_a = np.array([0.05, -0.1, 0.2]), np.array([0.01, 0.02, -0.03])
    dt = 1e-10
    f0, _ = _fluid_step(_wall_scene(n, 0.9 * _WALL_DX, v_i, v_a, 1.2, 1e-3, (0, 0, 0))[0], dt)
    f1, _ = _fluid_step(_wall_scene(n, 0.9 * _WALL_DX, v_i + u, v_a + u, 1.2, 1e-3, (0, 0, 0))[0], dt)
    r0, r1 = (f0["drho"][0] - 1.2) / dt, (f1["drho"][0] - 1.2) / dt
    assert r1 == pytest.approx(r0, rel=1e-6)
    assert np.abs(f1["acc"][0] - f0["acc"][0]).max() < 1e-7 * np.linalg.norm(f0["acc"][0])


def test_wall_riemann_speeds_use_the_wall_normal():
    # Off the normal axis, a velocity tangential to the wall has no normal
    # component: U_L = U_R, no Riemann dissipation, and the acceleration is
    # odd in that velocity around its value at rest (only the viscous term
    # moves).  Speeds taken along e_ia instead would switch the one-sided
    # dissipation on for one sign only.
    n = np.array([0.0, 0.0, 1.0])
    off = _unit((0.6, 0.0, 0.8))              # r_ia direction, not n
    t = np.array([1.0, 0.0, 0.0])             # tangential, with t.e_ia != 0
    dt = 1e-10

    def acc(v):
        sc, p_i = _wall_scene(n, 0.0, v, (0, 0, 0), 0.8, 1e-3, (0, 0, 0))
        sc.solid_pos = np.array([p_i - 0.9 * _WALL_DX * off])
        return _fluid_step(sc, dt)[0]["acc"][0]

    a0, ap, am = acc(0 * t), acc(0.3 * t), acc(-0.3 * t)
    assert np.linalg.norm(ap - a0) > 1e-6 * np.linalg.norm(a0)
    assert np.abs((ap - a0) + (am - a0)).max() < 1e-7 * np.linalg.norm(ap - a0)


def test_wall_hydrostatic_pressure_is_one_sided():
    # max(0, (g - a_a).n_a): gravity pointing out of the wall adds nothing
    n = _unit((0.1, 0.0, 1.0))
    v_i, v_a = (0.05, 0.0, -0.2), (0, 0, 0)
    dt = 1e-10
    f0, _ = _fluid_step(_wall_scene(n, 0.7 * _WALL_DX, v_i, v_a, 0.5, 1e-3, (0, 0, 0))[0], dt)
    g = np.array([0.0, 0.0, -9.81])           # g.n < 0
    f1, _ = _fluid_step(_wall_scene(n, 0.7 * _WALL_DX, v_i, v_a, 0.5, 1e-3, g)[0], dt)
    assert np.abs((f1["acc"][0] - g) - f0["acc"][0]).max() < 1e-7 * np.linalg.norm(f0["acc"][0])
    assert (f1["drho"][0] - 0.5) == pytest.approx(f0["drho"][0] - 0.5, rel=1e-6)
    f2, _ = _fluid_step(_wall_scene(n, 0.7 * _WALL_DX, v_i, v_a, 0.5, 1e-3, -g)[0], dt)
    assert np.linalg.norm((f2["acc"][0] + g) - f0["acc"][0]) > 1e-3 * np.linalg.norm(f0["acc"][0])


@pytest.mark.parametrize("g", [(0, 0, 0), (0, 0, -9.81), (0, 0, 9.81)])
def test_fluid_to_solid_force_is_the_reaction_of_the_wall_term(g):
    # calFluid2SolidForce (integrate.cu:282-330) is the wall term of
    # momentum times m_i, per pair: summed over the walls of a fluid
    # particle with no fluid neighbour it is m_i (a_i - g)
    sc = scenes.empty_scene((0, 0, 0), (0.1, 0.1, 0.1))
    sc.gravity = np.asarray(g, float)
    p_i = np.array([0.05, 0.05, 0.05])
    scenes.add_fluid(sc, [p_i], _H, _RHO0, _C, 1e-3)
    sc.fluid_vel = np.array([[0.05, -0.1, -0.2]])
    sc.fluid_drho = np.array([0.9])
    r_a = _H / 2.6
    grid = [(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1)]
    walls = [p_i + np.array([x * 2 * r_a, y * 2 * r_a, -0.7 * _WALL_DX]) for x, y in grid]
    scenes.add_dummy_wall(sc, walls, r_a)
    sc.solid_vel = np.tile([0.0, 0.02, 0.0], (len(walls), 1))
    sc.solid_normal = np.tile(_unit((0.05, 0.0, 1.0)), (len(walls), 1))
    f, fs = _fluid_step(sc, 1e-10)
    assert fs.shape == (len(walls), 3)
    m_i = (_H / 1.3) ** 3 * _RHO0
    total = fs.sum(axis=0)
    expect = m_i * (f["acc"][0] - np.asarray(g, float))
    assert np.linalg.norm(expect) > 0
    assert np.abs(total - expect).max() < 1e-9 * np.linalg.norm(expect)


def _wall_slab(nx, ny, nz, r):
    pts = np.array([[i, j, k] for k in range(nz) for j in range(ny) for i in range(nx)], float)
    return 0.02 + 2 * r * pts


def _normals(sc):
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize(compute_normals=True)
    nrm = o.solid()["normal"]
    o.close()
    return nrm


def test_wall_normals_point_out_of_a_slab():
    # setSolidNormal (dataBase.cpp:348-413): normalize(-sum V_B grad W) over
    # fixed neighbours within 2h, h the largest fluid h; a flat slab's top
    # layer points up, its bottom layer down, its middle layer (symmetric
    # neighbourhood) has no normal
    r = 0.0025
    sc = scenes.empty_scene((0, 0, 0), (0.2, 0.2, 0.2))
    scenes.add_fluid(sc, [[0.15, 0.15, 0.15]], 1.3 * 2 * r, _RHO0, _C, 1e-6)
    nx, nz = 15, 5
    pts = _wall_slab(nx, nx, nz, r)
    scenes.add_dummy_wall(sc, pts, r)
    nrm = _normals(sc)
    ijk = np.rint((pts - 0.02) / (2 * r)).astype(int)
    reach = 3   # 2h / spacing = 2.6: lattice neighbours within 3 spacings
    inner = np.all((ijk[:, :2] >= reach) & (ijk[:, :2] < nx - reach), axis=1)
    top = inner & (ijk[:, 2] == nz - 1)
    bot = inner & (ijk[:, 2] == 0)
    mid = inner & (ijk[:, 2] == nz // 2)
    assert top.sum() > 20
    assert np.abs(nrm[top] - [0, 0, 1]).max() < 1e-12
    assert np.abs(nrm[bot] - [0, 0, -1]).max() < 1e-12
    assert np.all(nrm[mid] == 0.0)


def test_wall_normals_bisect_the_edge_of_a_bar():
    # a bar of square section along y: an edge particle's neighbourhood is
    # symmetric under x <-> z, so its normal is (1, 0, 1)/sqrt(2) up to sign
    r = 0.0025
    sc = scenes.empty_scene((0, 0, 0), (0.2, 0.2, 0.2))
    scenes.add_fluid(sc, [[0.15, 0.15, 0.15]], 1.3 * 2 * r, _RHO0, _C, 1e-6)
    n = 6
    pts = _wall_slab(n, 16, n, r)
    scenes.add_dummy_wall(sc, pts, r)
    nrm = _normals(sc)
    ijk = np.rint((pts - 0.02) / (2 * r)).astype(int)
    along = (ijk[:, 1] >= 3) & (ijk[:, 1] < 13)
    for (ex, ez), sgn in (((n - 1, n - 1), (1, 1)), ((0, 0), (-1, -1)), ((n - 1, 0), (1, -1))):
        sel = along & (ijk[:, 0] == ex) & (ijk[:, 2] == ez)
        assert sel.sum() == 10
        expect = np.array([sgn[0], 0, sgn[1]]) / math.sqrt(2)
        assert np.abs(nrm[sel] - expect).max() < 1e-12


def test_density_reinit_counts_a_wall_particle_as_its_volume_of_fluid():
    # densityReinitialization (integrate.cu:3-53) adds W(r, h_i) (2 r_a)^3 / V0_i
    # per wall particle, V0_i = (h_i/1.3)^3: walls of radius h/2.6 on the
    # fluid lattice weigh exactly as the fluid particles they replace
    dx = 0.01
    h = 1.3 * dx
    fluid = scenes.regular_packed_points((0.03, 0.03, 0.03), (6 * dx,) * 3, dx)
    floor = scenes.regular_packed_points((0.03, 0.03, 0.0), (6 * dx, 6 * dx, 3 * dx), dx)

    def sumw0(walls_as_fluid):
        sc = scenes.empty_scene((-0.02, -0.02, -0.02), (0.14, 0.14, 0.14))
        sc.gravity = np.zeros(3)
        scenes.add_fluid(sc, fluid, h, _RHO0, _C, 1e-6)
        if walls_as_fluid:
            scenes.add_fluid(sc, floor, h, _RHO0, _C, 1e-6)
        else:
            scenes.add_dummy_wall(sc, floor, h / 2.6)
        o = oracle.Oracle()
        o.upload(sc)
        o.initialize(compute_normals=False)
        o.fluid_integrate(1e-9, 1, 1)    # step 1 stores sumW0
        out = o.fluid()["sumw0"][:len(fluid)]
        o.close()
        return out

    a, b = sumw0(False), sumw0(True)
    near = fluid[:, 2] < 0.03 + 2 * h
    assert near.sum() > 30
    assert np.abs(a - b).max() < 1e-12 * b.max()


def test_linear_contact_restitution_closed_form():
    # LinearContact (SPHKernelDEMContactModel.h:214-268): F_n = k_n d n -
    # 2 zeta sqrt(m* k_n) v_n, a damped oscillator whose contact ends at
    # d = 0 after half a damped period: e = exp(-pi zeta / sqrt(1 - zeta^2))
    kn, zeta = 1.0e5, 0.2
    params = [kn, 0.0, 0.0, 0.0, zeta, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    n = np.array([0, 0, 1.0])
    F, _, _ = oracle.contact(1, [0, 0, 0], [0, 0, 0], n, 3e-5, 0.01, 0.005, 1e-5, params)
    assert F[2] == pytest.approx(kn * 3e-5, rel=1e-14)
    e = _head_on_rebound(params, 0.5, 0.01, 2500.0, 1e-7, model=1)
    assert e == pytest.approx(math.exp(-math.pi * zeta / math.sqrt(1 - zeta * zeta)), abs=2e-3)


# --------------------------------------------- DEM sums and list history
def _forces_now(sc, dt=1e-6):
    """Solid accelerations and torques of the state as loaded, with the
    positions they were computed at (one search, a half kick-drift, the
    contact phase)."""
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize()
    o.neighbor_search(1, 1)
    o.solid_before_contact(dt)
    o.contact_forces(dt)
    s = o.solid()
    o.close()
    return s


def test_dem_contact_sums_conserve_momentum_and_angular_momentum():
    # calSolidForceTorque (integrate.cu:806-858): +F and +T + (r_c - r_i) x F
    # on the pointed solid, the negatives on the pointing one, both about the
    # same contact point, so sum F = 0 and sum (T + r x F) = 0 without
    # gravity or walls
    rng = np.random.default_rng(5)
    r = 0.004
    sc = scenes.empty_scene((-0.05, -0.05, -0.05), (0.1, 0.1, 0.1))
    sc.gravity = np.zeros(3)
    g = np.array([[i, j, k] for i in range(4) for j in range(4) for k in range(3)], float)
    pos = 1.9 * r * g + rng.uniform(-0.1 * r, 0.1 * r, g.shape)
    vel = rng.normal(scale=0.05, size=g.shape)
    scenes.add_solid(sc, pos, r, 2500.0, vel=vel)
    sc.solid_angvel = rng.normal(scale=5.0, size=g.shape)
    sc.set_hertz(0, 0, 5.4945e6, 1.131e6, 0.5, 0.3, 0.2, 0.4, 0.1, 0.1)
    s = _forces_now(sc)
    m = 1.0 / np.asarray(sc.solid_inv_mass)
    F = m[:, None] * s["acc"]
    scale = np.abs(F).sum()
    assert scale > 0 and np.abs(s["torque"]).sum() > 0
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * scale
    L = s["torque"] + np.cross(s["pos"], F)
    assert np.abs(L.sum(axis=0)).max() < 1e-12 * (np.abs(s["torque"]).sum() + np.abs(np.cross(s["pos"], F)).sum())


def test_fluid_force_on_a_sphere_is_the_reaction_of_the_fluid():
    # the fluid-solid pair forces enter the solid sum with a minus sign
    # (integrate.cu:846-853): fluid plus sphere exchange no momentum
    sc = scenes.empty_scene((0, 0, 0), (0.12, 0.12, 0.12))
    sc.gravity = np.zeros(3)
    dx = 0.005
    c = np.array([0.06, 0.06, 0.06])
    rs = 0.01
    pts = scenes.regular_packed_points((0.03, 0.03, 0.03), (0.06, 0.06, 0.06), dx)
    d = np.linalg.norm(pts - c, axis=1)
    pts = pts[(d > rs + 0.3 * dx) & (d < rs + 3 * dx)]
    scenes.add_fluid(sc, pts, 1.3 * dx, _RHO0, _C, 1e-3)
    rng = np.random.default_rng(9)
    sc.fluid_vel = rng.normal(scale=0.05, size=pts.shape)
    sc.fluid_drho = rng.normal(scale=0.5, size=len(pts))
    scenes.add_solid(sc, [c], rs, 1200.0, vel=(0.02, -0.01, 0.0))
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize()
    dt = 1e-7
    o.neighbor_search(2, 1)
    o.solid_before_contact(dt)
    o.fluid_integrate(dt, 2, 1)
    o.contact_forces(dt)
    fl, so, fs = o.fluid(), o.solid(), o.fs_force()
    o.close()
    assert len(fs) > 20
    m_s = 1.0 / float(sc.solid_inv_mass[0])
    m_f = (1.3 * dx / 1.3) ** 3 * _RHO0
    total_fs = fs.sum(axis=0)
    assert np.linalg.norm(total_fs) > 0
    assert np.abs(m_s * so["acc"][0] + total_fs).max() < 1e-12 * np.abs(fs).sum()
    P = m_f * fl["acc"].sum(axis=0) + m_s * so["acc"][0]
    assert np.abs(P).max() < 1e-10 * m_f * np.abs(fl["acc"]).sum()


def test_sliding_spring_history_carries_over_between_steps():
    # interactionSolid2Solid::save and the history copy
    # (neighborSearch.cu:287-301): a persisting contact keeps its springs,
    # so a steady tangential slip accumulates v_t dt per step
    r = 0.005
    sc = scenes.empty_scene((-0.05, -0.05, -0.05), (0.1, 0.1, 0.1))
    sc.gravity = np.zeros(3)
    v = 0.01
    scenes.add_solid(sc, [[0, 0, 0], [0, 0, 2 * r - 1e-5]], r, 1.0e9, vel=(v, 0, 0))
    sc.solid_inv_mass[0] = 0.0                                    # fixed
    sc.solid_vel[0] = 0.0
    sc.set_hertz(0, 0, 5.4945e6, 1.131e6, 1.0, 0, 0, 1e9, 0, 0)    # no damping, no cap
    o = _oracle_for(sc)
    dt = 1e-6
    k = 6
    for s in range(1, k + 1):
        o.step(s, dt)
    sl = o.contacts()["sliding"]
    o.close()
    assert sl.shape == (1, 3)
    assert np.linalg.norm(sl[0]) == pytest.approx(k * v * dt, rel=1e-3)


def test_bonds_join_touching_spheres_of_one_cluster_only():
    # addBondData (dataBase.cpp:270-346): a contact becomes a bond when the
    # spheres touch physically (r_i + r_j + 1e-10 > d), share a cluster and
    # their material pair has E_b > 1e-10
    r = 0.005
    sc = scenes.empty_scene((-0.05, -0.05, -0.05), (0.1, 0.1, 0.1))
    sc.gravity = np.zeros(3)
    scenes.add_cluster(sc, [[0, 0, 0], [2 * r, 0, 0], [4 * r, 0, 0], [6 * r + 1e-6, 0, 0]],
                       r, 2500.0)
    scenes.add_cluster(sc, [[-2 * r, 0, 0]], r, 2500.0)           # touches solid 0
    sc.set_hertz(0, 0, 5.4945e6, 1.131e6, 0.5, 0, 0, 0.4, 0, 0)
    sc.set_bonded(0, 0, 1e8, 2.0, 1.0, 1e6, 1e6, 0.3)
    o = _oracle_for(sc)
    a, b, _ = o.list(3)
    cand, _, _ = o.list(2)
    o.close()
    assert len(cand) >= 4                   # every pair above is a contact candidate
    assert sorted(zip(a.tolist(), b.tolist())) == [(0, 1), (1, 2)]


# ------------------------------------------------------------------ clumps
def _tetra_clump(sc, centroid, r, vel=(0, 0, 0), angvel=(0, 0, 0)):
    off = scenes.tetra_offsets(2 * r)
    m1 = 4 / 3 * math.pi * r ** 3 * 2500
    I = 4 * 0.4 * m1 * r * r + 4 * m1 * float(np.sum(off[0] ** 2)) * 2 / 3
    scenes.add_clump(sc, np.asarray(centroid) + off, np.full(4, r), np.asarray(centroid, float),
                     4 * m1, (I, I, 1.3 * I, 0, 0, 0), vel=vel, angvel=angvel)
    return 4 * m1


def test_clump_pebbles_move_rigidly():
    # clumpVelocityAngularVelocityIntegrate (integrate.cu:908-929): after the
    # clump kick every pebble has v_c + w x (r_p - r_c) and w
    sc = scenes.empty_scene((-0.1, -0.1, -0.1), (0.2, 0.2, 0.2))
    sc.gravity = np.array([0.0, 0.0, -9.81])
    _tetra_clump(sc, (0.01, -0.02, 0.03), 0.002, vel=(0.1, 0.0, -0.05), angvel=(3.0, -1.0, 2.0))
    o = _oracle_for(sc)
    for s in range(1, 21):
        o.step(s, 1e-4)
    c, p = o.clump(), o.solid()
    o.close()
    rel = p["pos"] - c["centroid"][0]
    expect = c["vel"][0] + np.cross(c["angvel"][0], rel)
    assert np.abs(p["vel"] - expect).max() < 1e-14
    assert np.abs(p["angvel"] - c["angvel"][0]).max() == 0.0
    # half kicks around the contact phase; the first one has no acceleration
    # yet, so 19.5 steps of gravity
    assert c["vel"][0, 2] == pytest.approx(-0.05 - 9.81 * 19.5 * 1e-4, rel=1e-12)


def test_clump_force_and_torque_sum_the_pebble_contact():
    # calClumpForceTorque (integrate.cu:860-881): the clump takes its
    # pebbles' forces and their torques about the centroid.  One pebble
    # touches a fixed sphere of its own radius, so the contact point is the
    # midpoint of the two centres.
    r = 0.002
    sc = scenes.empty_scene((-0.1, -0.1, -0.1), (0.2, 0.2, 0.2))
    sc.gravity = np.zeros(3)
    ctr = np.array([0.0, 0.0, 0.0])
    M = _tetra_clump(sc, ctr, r, vel=(0.02, 0.0, -0.1), angvel=(5.0, 2.0, -3.0))
    pebbles = np.asarray(sc.solid_pos)[:4]
    low = int(np.argmin(pebbles[:, 2]))
    p_low = pebbles[low]
    fixed = p_low - np.array([0.0, 0.3 * r, 1.99 * r])
    fixed = p_low + (fixed - p_low) / np.linalg.norm(fixed - p_low) * (2 * r - 2e-6)
    scenes.add_solid(sc, [fixed], r, 2500.0)
    sc.solid_inv_mass[-1] = 0.0
    sc.set_hertz(0, 0, 5.4945e6, 1.131e6, 0.5, 0.3, 0.2, 0.4, 0.1, 0.1)
    o = oracle.Oracle()
    o.upload(sc)
    o.initialize()
    o.neighbor_search(1, 1)
    dt = 1e-7
    o.solid_before_contact(dt)
    o.contact_forces(dt)
    c, so, ct = o.clump(), o.solid(), o.contacts()
    a, b, _ = o.list(2)
    o.close()
    assert len(a) == 1 and {int(a[0]), int(b[0])} == {low, 4}
    sgn = 1.0 if int(a[0]) == low else -1.0     # the force is on the pointed solid
    F, T = sgn * ct["force"][0], sgn * ct["torque"][0]
    mid = 0.5 * (so["pos"][low] + so["pos"][4])
    assert np.linalg.norm(F) > 0 and np.linalg.norm(T) > 0
    assert np.abs(M * c["acc"][0] - F).max() < 1e-12 * np.linalg.norm(F)
    Tc = T + np.cross(mid - c["centroid"][0], F)
    assert np.abs(c["torque"][0] - Tc).max() < 1e-12 * np.linalg.norm(Tc)


# ------------------------------------- runtime mutation (dataBase.cpp:179-224)
def _lone_sphere(v=(0.0, 0.0, 0.0), g=(0.0, 0.0, 0.0)):
    sc = scenes.empty_scene((-1, -1, -1), (2, 2, 2))
    sc.gravity = np.asarray(g, float)
    scenes.add_solid(sc, [[0.0, 0.0, 0.0]], 0.01, 2500.0, material=0, vel=v)
    return sc


def _step_with_hook(o, s, dt, hook):
    o.neighbor_search(s, 1)
    o.solid_before_contact(dt)
    o.fluid_integrate(dt, s, 1)
    o.contact_forces(dt)
    hook(o)
    o.solid_after_contact(dt)


def test_external_force_accelerates_a_free_sphere():
    # a constant force F in the user hook of every step: the half kicks
    # around it give v = (N - 1/2) F dt / m after N steps (the first
    # step's opening half kick sees no force yet)
    sc = _lone_sphere()
    m = 1.0 / float(sc.solid_inv_mass[0])
    F = np.array([2e-3, -1e-3, 5e-4])
    dt, n = 1e-4, 25
    o = _oracle_for(sc)
    for s in range(1, n + 1):
        _step_with_hook(o, s, dt, lambda x: x.add_external_force(0, F))
    v = o.solid()["vel"][0]
    assert np.allclose(v, (n - 0.5) * dt * F / m, rtol=1e-12, atol=0)


def test_global_damping_opposes_the_velocity():
    # gravity only, moving along x: F = m g, |F| = m |g|, and the damping
    # takes C_d |F| along -v, a deceleration C_d |g| along x; one step
    # (the opening half kick sees no acceleration yet)
    g = 9.81
    cd, dt = 0.2, 1e-3
    sc = _lone_sphere(v=(1.0, 0.0, 0.0), g=(0, 0, -g))
    o = _oracle_for(sc)
    _step_with_hook(o, 1, dt, lambda x: x.add_global_damping(0, cd))
    v = o.solid()["vel"][0]
    assert v[0] == pytest.approx(1.0 - 0.5 * dt * cd * g, rel=1e-14)
    assert v[2] == pytest.approx(-0.5 * dt * g, rel=1e-14)
    # a fixed solid is left alone, and so is one at rest
    sc2 = _lone_sphere(g=(0, 0, -g))
    o2 = _oracle_for(sc2)
    _step_with_hook(o2, 1, dt, lambda x: x.add_global_damping(0, cd))
    assert o2.solid()["vel"][0][0] == 0.0


# ------------------------------------------------ cylinder walls (K16)
def _settle_against_cylinder(inside, g_vec, pos, axis, R):
    E, G = 5.4945e6, 1.131e6
    r, rho = 0.01, 2500.0
    sc = scenes.empty_scene((-0.2, -0.2, -0.2), (0.4, 0.4, 0.4))
    sc.gravity = np.asarray(g_vec, float)
    scenes.add_solid(sc, [pos], r, rho)
    sc.set_hertz(0, 0, E, G, 0.05, 0, 0, 0.4, 0, 0)
    scenes.add_cylinder_wall(sc, (0, 0, 0), axis, (0, 0, 0), R, 0)
    o = _oracle_for(sc)
    for s in range(1, 6001):
        o.step(s, 2e-5)
    m = 4 / 3 * math.pi * r ** 3 * rho
    d_expect = (m * np.linalg.norm(g_vec) / (4 / 3 * E * math.sqrt(r))) ** (2 / 3)
    return o.solid()["pos"][0], r, d_expect


def test_sphere_resting_inside_a_pipe_static_overlap():
    # a horizontal pipe (axis x) of radius R holds a sphere at its bottom:
    # the contact law takes the sphere's radius, so the overlap is that of
    # a sphere on a plane, and the sphere sits R - r + delta below the axis
    R, r = 0.05, 0.01
    p, r, d = _settle_against_cylinder(True, (0, 0, -9.81), (0.0, 0.0, -(R - r) * 0.999), (2, 0, 0), R)
    assert abs(p[0]) < 1e-12 and abs(p[1]) < 1e-12
    assert -p[2] - (R - r) == pytest.approx(d, rel=2e-3)


def test_sphere_pressed_on_the_outside_of_a_cylinder():
    # gravity along -x presses a sphere against the outside of a vertical
    # cylinder (axis z) at x = R + r: the overlap is the plane's again
    R, r = 0.03, 0.01
    p, r, d = _settle_against_cylinder(False, (-9.81, 0, 0), ((R + r) * 1.001, 0.0, 0.0), (0, 0, 1), R)
    assert R + r - p[0] == pytest.approx(d, rel=2e-3)
    assert abs(p[1]) < 1e-12 and abs(p[2]) < 1e-12


def test_rotating_pipe_drags_a_resting_sphere_along_its_wall():
    # the pipe turns about its axis: friction at the contact pushes the
    # sphere the way the wall moves there (and spins it the other way)
    R, r, rho = 0.05, 0.01, 2500.0
    sc = scenes.empty_scene((-0.2, -0.2, -0.2), (0.4, 0.4, 0.4))
    scenes.add_solid(sc, [[0.0, 0.0, -(R - r)]], r, rho)
    sc.set_hertz(0, 0, 5.4945e6, 1.131e6, 0.5, 0, 0, 0.4, 0, 0)
    w = np.array([5.0, 0.0, 0.0])                     # about the pipe axis x
    scenes.add_cylinder_wall(sc, (0, 0, 0), (1, 0, 0), w, R, 0)
    o = _oracle_for(sc)
    for s in range(1, 201):
        o.step(s, 2e-5)
    v = o.solid()["vel"][0]
    wall_v = np.cross(w, [0.0, 0.0, -R])               # the wall's velocity under the sphere
    assert wall_v[1] > 0 and v[1] > 0
    assert o.solid()["angvel"][0][0] > 0               # rolls with the wall's sense of turn
    # a pipe at rest leaves it at rest sideways
    sc2 = scenes.empty_scene((-0.2, -0.2, -0.2), (0.4, 0.4, 0.4))
    scenes.add_solid(sc2, [[0.0, 0.0, -(R - r)]], r, rho)
    sc2.hertz = sc.hertz
    scenes.add_cylinder_wall(sc2, (0, 0, 0), (1, 0, 0), (0, 0, 0), R, 0)
    o2 = _oracle_for(sc2)
    for s in range(1, 201):
        o2.step(s, 2e-5)
    assert abs(o2.solid()["vel"][0][1]) < 1e-12
