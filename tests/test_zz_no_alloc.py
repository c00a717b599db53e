"""dsph_hash_sort_noalloc (the allocation-free variant of dsph_hash_sort,
whose work buffer is a fixed device array of the library) against
dsph_hash_sort through the C ABI: bit-identical hashes, permutations and
cell ranges on two calls in a row, the second with a smaller input, and
the size limit rejected with a clear error."""
import numpy as np
import pytest

import dsph

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_POINTS = 1 << 20


def call(g, pos, origin, size, cells, noalloc):
    n = len(pos)
    ncell = int(np.prod(cells)) + 1
    dev = torch.device("cuda")
    p = torch.from_numpy(np.ascontiguousarray(pos, np.float64)).to(dev)
    out = [torch.full((max(n, 1),), -7, dtype=torch.int32, device=dev) for _ in range(2)]
    out += [torch.full((ncell,), -7, dtype=torch.int32, device=dev) for _ in range(2)]
    torch.cuda.synchronize()
    g.hash_sort(p.data_ptr(), n, origin, size, cells, *(t.data_ptr() for t in out), noalloc=noalloc)
    g.synchronize()
    h, pm, s, e = (t.cpu().numpy() for t in out)
    return h[:n], pm[:n], s, e


def points(n, origin, size, seed):
    rng = np.random.default_rng(seed)
    lo = np.asarray(origin) - 0.05
    hi = np.asarray(origin) + np.asarray(size) + 0.05
    return rng.uniform(lo, hi, (n, 3))


def test_noalloc_matches_hash_sort_on_two_calls():
    g = dsph.Solver(dsph.PREC_FP64)
    # a large call, then a smaller one on another grid: the second must not
    # see anything of the first in the shared scratch
    cases = [(MAX_POINTS, (-0.3, 0.1, -1.0), (1.7, 0.6, 0.9), (61, 23, 33)),
             (40_000, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (7, 5, 3))]
    for k, (n, origin, size, cells) in enumerate(cases):
        pos = points(n, origin, size, 100 + k)
        a = call(g, pos, origin, size, cells, noalloc=False)
        b = call(g, pos, origin, size, cells, noalloc=True)
        for x, y, name in zip(a, b, ("hash", "perm", "start", "end")):
            assert np.array_equal(x, y), (k, name)
        # and the result is a stable sort of the hashes into cell ranges
        h, perm, start, end = b
        assert np.array_equal(perm, np.argsort(h, kind="stable"))
        occupied = np.unique(h)
        assert np.array_equal(start[occupied], np.searchsorted(h[perm], occupied, "left"))
        assert np.array_equal(end[occupied], np.searchsorted(h[perm], occupied, "right"))
        assert np.all(start[np.setdiff1d(np.arange(len(start)), occupied)] == -1)


def test_noalloc_rejects_more_points_than_its_scratch():
    g = dsph.Solver(dsph.PREC_FP64)
    with pytest.raises(dsph.DsphError) as e:
        # the size check comes before any device access
        g.hash_sort(1, MAX_POINTS + 1, (0, 0, 0), (1, 1, 1), (2, 2, 2), 1, 1, 1, 1, noalloc=True)
    assert e.value.code == -1 and "1048576" in str(e.value)
