"""dsph_hash_sort: calculatePointHash, the stable hash sort and the cell
ranges (neighborSearch.cu:19-113) on caller device arrays, against numpy
(refmodel): truncating hash with a spare cell, a stable argsort, ranges."""
import numpy as np
import pytest

import dsph
import refmodel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def reference(pos, origin, size, cells):
    G = np.asarray(cells)
    cell = np.asarray(size, float) / G
    ncell = int(np.prod(G)) + 1
    h, _ = refmodel.point_hash(pos, np.asarray(origin, float), np.asarray(size, float), G, cell, ncell)
    perm = np.argsort(h, kind="stable")
    start, end = refmodel.cell_ranges(h[perm], ncell)
    return h, perm, start, end, ncell


def run(g, pos, origin, size, cells, ncell):
    n = len(pos)
    dev = torch.device("cuda")
    p = torch.from_numpy(np.ascontiguousarray(pos, np.float64)).to(dev)
    h = torch.full((max(n, 1),), -7, dtype=torch.int32, device=dev)
    pm = torch.full((max(n, 1),), -7, dtype=torch.int32, device=dev)
    s = torch.full((ncell,), -7, dtype=torch.int32, device=dev)
    e = torch.full((ncell,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    g.hash_sort(p.data_ptr(), n, origin, size, cells, h.data_ptr(), pm.data_ptr(), s.data_ptr(), e.data_ptr())
    g.synchronize()
    return h.cpu().numpy()[:n], pm.cpu().numpy()[:n], s.cpu().numpy(), e.cpu().numpy()


@pytest.mark.parametrize("n,cells", [(200_000, (37, 23, 11)), (5000, (1, 1, 1)), (1, (4, 4, 4)),
                                     (0, (3, 5, 7)), (70_000, (96, 2, 64))])
def test_hash_sort_matches_numpy(n, cells):
    rng = np.random.default_rng(n + 11)
    origin, size = (-0.3, 0.1, -1.0), (1.7, 0.6, 0.9)
    lo = np.asarray(origin) - 0.05
    hi = np.asarray(origin) + np.asarray(size) + 0.05   # a few points outside: spare cell
    pos = rng.uniform(lo, hi, (n, 3))
    h, perm, start, end, ncell = reference(pos, origin, size, cells)
    g = dsph.Solver(dsph.PREC_FP64)
    gh, gp, gs, ge = run(g, pos, origin, size, cells, ncell)
    assert np.array_equal(gh, h) and np.array_equal(gp, perm)
    assert np.array_equal(gs, start) and np.array_equal(ge, end)


def test_hash_sort_rejects_bad_grids():
    g = dsph.Solver(dsph.PREC_FP64)
    with pytest.raises(dsph.DsphError):
        g.hash_sort(0, 0, (0, 0, 0), (1, 1, 1), (0, 2, 2), 0, 0, 1, 1)
    with pytest.raises(dsph.DsphError):
        g.hash_sort(0, 0, (0, 0, 0), (1, -1, 1), (2, 2, 2), 0, 0, 1, 1)
