"""Independent brute-force / closed-form checkers used to pin the oracle.

These are written from the mathematics of the reference's definitions, not by
re-typing the oracle: O(N^2) neighbour sets, the stencil-order rule for list
order, stable-sort permutations via numpy, quadrature of the SPH kernel.
"""
from __future__ import annotations

import numpy as np


def grid_params(scene):
    """setSpatialGrids (core/dataBase.cpp:415-432)."""
    mf = float(np.max(scene.fluid_h)) if scene.n_fluid else 0.0
    ms = float(np.max(scene.solid_eff_radius)) if scene.n_solid else 0.0
    cs = 2 * max(mf, ms)
    size = np.asarray(scene.domain_size, float)
    G = [int(s / cs) if s > cs else 1 for s in size]
    cell = [float(size[k]) / float(G[k]) for k in range(3)]
    ncell = G[0] * G[1] * G[2] + 1
    return np.asarray(scene.domain_origin, float), size, np.array(G), np.array(cell), ncell


def point_hash(pos, origin, size, G, cell, ncell):
    """calculatePointHash: per axis int((x-min)/cell) truncation; the spare
    cell ncell-1 for points outside [min, min+size)."""
    pos = np.asarray(pos, float).reshape(-1, 3)
    mx = origin + size
    inside = np.all((origin <= pos) & (pos < mx), axis=1)
    g = np.trunc((pos - origin) / cell).astype(np.int64)
    h = g[:, 2] * G[1] * G[0] + g[:, 1] * G[0] + g[:, 0]
    return np.where(inside, h, ncell - 1).astype(np.int64), g


def cell_ranges(sorted_keys, ncell):
    start = np.full(ncell, -1, np.int64)
    end = np.full(ncell, -1, np.int64)
    n = len(sorted_keys)
    for i in range(n):
        k = sorted_keys[i]
        if i == 0 or sorted_keys[i - 1] != k:
            start[k] = i
        if i == n - 1 or sorted_keys[i + 1] != k:
            end[k] = i + 1
    return start, end


def stencil_rank(ga, gb, G):
    """Rank (0..26, z-major) of b's cell in a's stencil, from grid coords."""
    d = gb - ga
    return (d[..., 2] + 1) * 9 + (d[..., 1] + 1) * 3 + (d[..., 0] + 1)


def brute_fluid_lists(scene):
    """Reference-order fluid–fluid (j>i, |r| <= 2max(h)) and fluid–solid
    (|r| <= 2h_i) lists by O(N^2) search; order = stencil rank, then index.
    Valid for in-domain particles in grids with >= 3 cells per axis."""
    origin, size, G, cell, ncell = grid_params(scene)
    fp = np.asarray(scene.fluid_pos, float).reshape(-1, 3)
    h = np.asarray(scene.fluid_h, float)
    _, gf = point_hash(fp, origin, size, G, cell, ncell)
    sp = np.asarray(scene.solid_pos, float).reshape(-1, 3)
    _, gs = point_hash(sp, origin, size, G, cell, ncell)
    ff_i, ff_j, fs_i, fs_a = [], [], [], []
    for i in range(len(fp)):
        d = fp - fp[i]
        dis = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        rad = np.maximum(h[i], h)
        ok = (2 * rad - dis >= 0) & (np.arange(len(fp)) > i)
        js = np.nonzero(ok)[0]
        rk = stencil_rank(gf[i], gf[js], G)
        order = np.lexsort((js, rk))
        ff_i += [i] * len(js)
        ff_j += list(js[order])
        if len(sp):
            d = sp - fp[i]
            dis = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
            ok = 2 * h[i] - dis >= 0
            a = np.nonzero(ok)[0]
            rk = stencil_rank(gf[i], gs[a], G)
            order = np.lexsort((a, rk))
            fs_i += [i] * len(a)
            fs_a += list(a[order])
    return (np.array(ff_i, np.int64), np.array(ff_j, np.int64),
            np.array(fs_i, np.int64), np.array(fs_a, np.int64))


def brute_solid_list(scene):
    """DEM contact list (setSolid2SolidInteractionsKernel filters) by brute force."""
    origin, size, G, cell, ncell = grid_params(scene)
    sp = np.asarray(scene.solid_pos, float).reshape(-1, 3)
    eff = np.asarray(scene.solid_eff_radius, float)
    dummy = np.asarray(scene.solid_dummy, int)
    clump = np.asarray(scene.solid_clump, int)
    invm = np.asarray(scene.solid_inv_mass, float)
    _, gs = point_hash(sp, origin, size, G, cell, ncell)
    I, J = [], []
    idx = np.arange(len(sp))
    for a in range(len(sp)):
        if dummy[a] == 1:
            continue
        d = sp - sp[a]
        dis = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        ok = (dummy != 1) & (idx > a) & ~((clump[a] >= 0) & (clump == clump[a]))
        ok &= ~((invm[a] < 1e-10) & (invm < 1e-10))
        ok &= (eff + eff[a]) - dis >= 0
        b = np.nonzero(ok)[0]
        rk = stencil_rank(gs[a], gs[b], G)
        order = np.lexsort((b, rk))
        I += [a] * len(b)
        J += list(b[order])
    return np.array(I, np.int64), np.array(J, np.int64)


def wendland(r, h):
    q = np.asarray(r, float) / h
    sigma = 21.0 / (16.0 * np.pi * h ** 3)
    return np.where(q < 2, sigma * (1 - 0.5 * q) ** 4 * (1 + 2 * q), 0.0)


def rel_err(a, b, floor=0.0):
    """Norm-wise relative error max|a-b| / max(max|b|, floor)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    den = max(float(np.max(np.abs(b))) if b.size else 0.0, floor)
    if den == 0:
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a - b))) / den
