"""svdag_query_voxels_noalloc (staging through the library's static per-GPU
array, no device allocation) against svdag_query_voxels on the same DAG:
identical materials on two calls in a row, the second smaller, on two
contexts of one GPU, plus the oracle on the first call and the size limit."""
import numpy as np
import pytest

import oracle_py
import svdag_py
import workloads


@pytest.mark.gpu
def test_noalloc_query_equals_allocating_query():
    levels = 10
    tp = workloads.terrain_params(levels)
    g = svdag_py.MyGPUHashDAG(levels, 4)
    o = oracle_py.OracleDag(levels, 4)
    g.build_terrain(**tp)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    rng = np.random.default_rng(11)
    big = rng.integers(0, 1 << levels, size=(svdag_py.QUERY_NOALLOC_MAX, 3)).astype(np.uint32)
    small = rng.integers(0, 1 << levels, size=(4097, 3)).astype(np.uint32)
    a = g.query(big, noalloc=True)
    assert np.array_equal(a, g.query(big))
    assert np.array_equal(a[:200000], o.query(big[:200000]))
    assert (a != svdag_py.NOT_FOUND).any() and (a == svdag_py.NOT_FOUND).any()
    # Second call, smaller: the staging array still holds the first call's
    # points and results past n, which must not leak into this one.
    b = g.query(small, noalloc=True)
    assert len(b) == len(small)
    assert np.array_equal(b, g.query(small))

    # A second context on the same GPU shares the array.
    g2 = svdag_py.MyGPUHashDAG(8, 4)
    g2.build_terrain(**workloads.terrain_params(8))
    pts = rng.integers(0, 1 << 8, size=(30000, 3)).astype(np.uint32)
    assert np.array_equal(g2.query(pts, noalloc=True), g2.query(pts))
    assert np.array_equal(g.query(small, noalloc=True), b)

    assert len(g.query(np.zeros((0, 3), np.uint32), noalloc=True)) == 0
    with pytest.raises(svdag_py.SvdagError):
        g.query(np.zeros((svdag_py.QUERY_NOALLOC_MAX + 1, 3), np.uint32), noalloc=True)
    g2.close()
    g.close()
