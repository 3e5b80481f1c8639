"""The CPU HashDAG restatement (the CPU timing baseline) holds the right
voxels: an imported canonical DAG reads back voxel for voxel, and its
sphere / box edits, threaded or not, equal a brute-force voxelisation of the
reference CPU tools (hash_dag_editors.h: sphere without the GPU tools' +0.5
centre, box [c - r, c + r]; a voxel is edited iff its unit box passes the
tool's overlap test)."""
import numpy as np
import pytest

import oracle_py
import workloads


def _grid(levels):
    side = 1 << levels
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    return g.astype(np.uint32)


def _overlap(kind, c, r, pts):
    f = np.float32
    mn = pts.astype(np.float32)
    mx = mn + f(1.0)
    c = np.array(c, np.float32)
    r = f(r)
    if kind in (0, 1):
        dist = np.full(len(pts), r * r, np.float32)
        for a in range(3):
            lo = c[a] < mn[:, a]
            hi = c[a] > mx[:, a]
            dist = np.where(lo, dist - (c[a] - mn[:, a]) * (c[a] - mn[:, a]), dist)
            dist = np.where(~lo & hi, dist - (c[a] - mx[:, a]) * (c[a] - mx[:, a]), dist)
        return dist > 0
    bmin, bmax = c - r, c + r
    ok = np.ones(len(pts), bool)
    for a in range(3):
        ok &= ~((bmin[a] >= mx[:, a]) | (bmax[a] <= mn[:, a]))
    return ok


EDITS = [(0, (40.0, 70.0, 50.0), 20.0), (2, (90.0, 30.0, 80.0), 12.0), (1, (45.0, 65.0, 52.0), 11.0),
         (3, (88.0, 28.0, 84.0), 6.0), (0, (100.5, 100.5, 20.5), 9.5), (2, (10.0, 10.0, 10.0), 0.5),
         (1, (64.0, 64.0, 64.0), 70.0), (0, (20.0, 20.0, 100.0), 3.0)]


@pytest.mark.parametrize("threads", [1, 8])
def test_cpu_edits_match_brute_force(threads):
    levels = 7
    d = oracle_py.CpuHashDag(levels, pool_bytes=1 << 30)
    pts = _grid(levels)
    vox = np.zeros(len(pts), bool)
    for kind, c, r in EDITS:
        _, tasks = d.edit(kind, c, r, threads=threads)
        ov = _overlap(kind, c, r, pts)
        vox = (vox | ov) if kind in (0, 2) else (vox & ~ov)
        assert tasks > 0
        assert np.array_equal(d.query(pts).astype(bool), vox), (kind, c, r)
    assert vox.any()


def test_import_of_a_canonical_terrain_then_edits():
    levels = 8
    tp = workloads.terrain_params(levels)
    o = oracle_py.OracleDag(levels, 4)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    _, _, dump = o.canonical(dump=True)
    d = oracle_py.CpuHashDag(levels, pool_bytes=1 << 30)
    d.import_canonical(dump, threads=4)
    rng = np.random.default_rng(3)
    pts = rng.integers(0, 1 << levels, size=(200000, 3)).astype(np.uint32)
    want = o.query(pts) != 0xFFFFFFFF
    assert np.array_equal(d.query(pts).astype(bool), want)
    words, pages = d.usage()
    assert words > 0 and pages > 0
    # Edits over the imported terrain, 32 threads, against brute force.
    for kind, c, r in [(1, (128.0, 150.0, 128.0), 40.0), (0, (60.0, 200.0, 60.0), 25.0), (2, (200.0, 90.0, 40.0), 16.0)]:
        d.edit(kind, c, r, threads=32)
        ov = _overlap(kind, c, r, pts)
        want = (want | ov) if kind in (0, 2) else (want & ~ov)
        assert np.array_equal(d.query(pts).astype(bool), want)


def test_pool_overflow_is_reported():
    # A pool of 64 pages cannot hold a depth-8 terrain (every bucket in use
    # maps a page of its own): the import fails loudly instead of writing
    # past the pool.
    levels = 8
    tp = workloads.terrain_params(levels)
    o = oracle_py.OracleDag(levels, 4)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    _, _, dump = o.canonical(dump=True)
    d = oracle_py.CpuHashDag(levels, pool_bytes=64 * 512 * 4)
    with pytest.raises(RuntimeError):
        d.import_canonical(dump)
