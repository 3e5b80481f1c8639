"""GPU editable DAG against the oracle (CPU restatement of the reference GPU
path): canonical DAG, per-level edit-SVO sizes, voxel values and traced hit
buffers, bit-exact."""
import numpy as np
import pytest

import oracle_py
import svdag_py
import workloads

L10 = 10


def _pair(levels=L10, material_bits=4, group_size=0, terrain=True, **kw):
    g = svdag_py.MyGPUHashDAG(levels, material_bits, group_size=group_size, **kw)
    o = oracle_py.OracleDag(levels, material_bits)
    if terrain:
        tp = workloads.terrain_params(levels)
        g.build_terrain(**tp)
        o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    return g, o


def _stored_items(g):
    # Distinct (table, pointer) pairs among the records reachable from the GC roots.
    rec = g.export_records()
    n, p, seen = int(rec[3]), 4, set()
    leaf_level = g.levels - 2
    leaf_bits = {0: 0, 4: 3, 8: 4}[g.material_bits]
    for _ in range(n):
        level, handle, size = int(rec[p]), int(rec[p + 1]), int(rec[p + 2])
        bits = leaf_bits if level == leaf_level else 3
        ptr = handle & ((1 << (32 - bits)) - 1) if bits else handle
        seen.add((size, ptr))
        p += 3 + size
    return len(seen)


def _same(g, o):
    gh, gc = g.canonical()
    oh, oc = o.canonical()
    assert gc == oc, (gc, oc)
    assert gh == oh


@pytest.mark.gpu
def test_terrain_heights_bit_exact():
    g = svdag_py.MyGPUHashDAG(L10, 4)
    tp = workloads.terrain_params(L10)
    g.build_terrain(**tp)
    gh = g.terrain_heights(0, 0, 1024, 1024)
    oh = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                   0, 0, 1024, 1024)
    assert np.array_equal(gh, oh)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("group_size", [4, 8, 16, 32])
def test_terrain_canonical_matches_factory_rule(group_size):
    g, o = _pair(group_size=group_size)
    _same(g, o)
    # The oracle's canonicaliser over the GPU's own reachable records agrees too.
    rh, rc = oracle_py.canonical_from_records(g.export_records())
    assert rh == g.canonical()[0]
    g.close()


@pytest.mark.gpu
def test_config1_edits_parity():
    g, o = _pair()
    for kind, mat, c, r in workloads.config1_edits():
        st = g.edit(kind, mat, c, r)
        applied, counts = o.edit(kind, mat, c, r)
        assert st.applied == applied
        assert list(st.svo_items)[:L10 - 1] == counts[:L10 - 1]
        _same(g, o)
    g.close()


@pytest.mark.gpu
def test_surface_edit_stream_all_tools():
    g, o = _pair()
    heights = oracle_py.terrain_heights(**{k: v for k, v in dict(
        seed=workloads.terrain_params(L10)["seed"], base=512.0, amp=200.0, freq=1 / 256, octaves=5,
        x0=0, z0=0, w=1024, d=1024).items()})
    edits = workloads.surface_edits(lambda x, z: heights[z, x], 24, seed=6, levels=L10, rmin=2.0, rmax=90.0,
                                    near=40.0)
    edits += [(svdag_py.BOX_REMOVE, 0, (500.0, 500.0, 500.0), 37.0), (svdag_py.PAINT, 9, (300.0, 520.0, 300.0), 60.0),
              (svdag_py.SPHERE_ADD, 3, (100.0, 600.0, 900.0), 0.0), (svdag_py.BOX_ADD, 4, (700.0, 450.0, 200.0), 0.0)]
    for i, (kind, mat, c, r) in enumerate(edits):
        st = g.edit(kind, mat, c, r)
        applied, counts = o.edit(kind, mat, c, r)
        assert st.applied == applied, i
        assert list(st.svo_items)[:L10 - 1] == counts[:L10 - 1], i
    _same(g, o)
    # Copy/paste a region (translation only, as MyGpuCopyEditor).
    for src, dst, c, r in [((300.0, 500.0, 300.0), (600.0, 520.0, 650.0), (310.0, 505.0, 305.0), 40.0),
                           ((100.0, 480.0, 700.0), (103.0, 470.0, 690.0), (100.0, 480.0, 700.0), 25.0)]:
        st = g.edit(svdag_py.COPY, 0, c, r, src, dst)
        applied, counts = o.edit(svdag_py.COPY, 0, c, r, src, dst)
        assert st.applied == applied
        assert list(st.svo_items)[:L10 - 1] == counts[:L10 - 1]
        _same(g, o)
    g.close()


@pytest.mark.gpu
def test_voxel_queries_match_oracle():
    g, o = _pair()
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    rng = np.random.default_rng(5)
    pts = rng.integers(0, 1024, size=(200000, 3)).astype(np.uint32)
    c = np.array(workloads.config1_edits()[0][2], np.float32).astype(np.int64)
    box = (c + rng.integers(-70, 71, size=(100000, 3))).clip(0, 1023).astype(np.uint32)
    pts = np.concatenate([pts, box])
    assert np.array_equal(g.query(pts), o.query(pts))
    g.close()


@pytest.mark.gpu
def test_undo_redo_and_gc():
    g, o = _pair()
    h0 = g.canonical()[0]
    edits = workloads.config1_edits()
    hs = [h0]
    for kind, mat, c, r in edits:
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
        hs.append(g.canonical()[0])
    assert g.history() == (2, 0)
    g.undo()
    assert g.canonical()[0] == hs[1]
    g.undo()
    assert g.canonical()[0] == hs[0]
    g.redo()
    g.redo()
    assert g.canonical()[0] == hs[2]
    before = _stored_items(g)
    st = g.garbage_collect()
    assert g.history() == (0, 0)
    assert g.canonical()[0] == hs[2] == o.canonical()[0]
    # Live = the distinct stored items reachable from the GC roots (root and
    # fully-filled chains); the export fails if any of them was swept.
    assert st.items_live == _stored_items(g)
    assert st.items_freed > 0
    assert st.items_freed + st.items_live >= before
    # Edits after a collection reuse freed slots and stay exact.
    g.edit(svdag_py.SPHERE_ADD, 5, (512.0, 520.0, 512.0), 30.0)
    o.edit(svdag_py.SPHERE_ADD, 5, (512.0, 520.0, 512.0), 30.0)
    _same(g, o)
    g.close()


@pytest.mark.gpu
def test_edit_without_leaves_is_skipped():
    # A box exactly covering one level-1 octant makes no leaf: the reference
    # skips the merge (my_gpu_hash_dag_edits.cu:651), so nothing changes.
    g, o = _pair()
    before = g.canonical()[0]
    st = g.edit(svdag_py.BOX_ADD, 7, (256.0, 256.0, 256.0), 256.0)
    applied, _ = o.edit(svdag_py.BOX_ADD, 7, (256.0, 256.0, 256.0), 256.0)
    assert st.applied == 0 and applied == 0
    assert g.canonical()[0] == before == o.canonical()[0]
    g.close()


@pytest.mark.gpu
def test_trace_512_bit_exact():
    g, o = _pair()
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    gh, ms = g.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 512, 512)
    oh = o.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 512, 512)
    assert np.array_equal(gh.view(np.uint32).reshape(-1, 8), oh.view(np.uint32).reshape(-1, 8))
    hit = gh["material"] != 0xFFFFFFFF
    assert hit.mean() > 0.3
    # Every hit voxel is set, with the material the DAG holds.
    xyz = np.stack([gh["x"][hit], gh["y"][hit], gh["z"][hit]], 1)
    assert np.array_equal(g.query(xyz), gh["material"][hit])
    g.close()


@pytest.mark.gpu
def test_paint_sphere_inside_box():
    # Port of "Paint a sphere inside a box" (tests/gpu_hash_dag/gpu_hash_dag_tests.cpp:198-262).
    g, o = _pair(terrain=False)
    lo, hi = 100, 130
    c = (lo + hi) / 2.0
    for d in (g, o):
        d.edit(svdag_py.BOX_ADD, 0, (c, c, c), (hi - lo) / 2.0)
    center = (c + 10.0, c, c)
    radius = (hi - lo) / 2.0
    g.edit(svdag_py.PAINT, 13, center, radius)
    o.edit(svdag_py.PAINT, 13, center, radius)
    _same(g, o)
    pts = np.array([(x, y, z) for x in range(lo - 1, hi + 2) for y in range(lo - 1, hi + 2)
                    for z in range(lo - 1, hi + 2)], np.uint32)
    vals = g.query(pts)
    inside = np.all((pts >= lo) & (pts < hi), axis=1)
    assert np.all(vals[~inside] == 0xFFFFFFFF)
    assert np.all(vals[inside] != 0xFFFFFFFF)
    g.close()


@pytest.mark.gpu
def test_fully_filled_gate_quirk_parity():
    # The GPU reproduces the reference's flagless full node (gate at
    # my_gpu_hash_dag_edits.cu:337) exactly as the oracle does.
    from test_oracle_dag import GATE_CASE
    c = GATE_CASE
    g = svdag_py.MyGPUHashDAG(c["levels"], 4)
    o = oracle_py.OracleDag(c["levels"], 4)
    seed, base, amp, freq, octaves, mat = c["terrain"]
    g.build_terrain(seed, base, amp, freq, octaves, mat)
    o.build_terrain(*c["terrain"])
    for e in c["edits"]:
        assert g.edit(*e).applied == o.edit(*e)[0] == 1
    _same(g, o)
    g.close()


@pytest.mark.gpu
def test_batched_edits_equal_sequential_edits():
    heights = oracle_py.terrain_heights(workloads.terrain_params(L10)["seed"], 512.0, 200.0, 1 / 256, 5, 0, 0, 1024, 1024)
    edits = workloads.surface_edits(lambda x, z: heights[z, x], 20, seed=9, levels=L10, rmin=2.0, rmax=120.0, near=60.0)
    edits.append((svdag_py.BOX_ADD, 7, (256.0, 256.0, 256.0), 256.0))  # makes no leaf: skipped in both forms
    edits.append((svdag_py.COPY, 0, (200.0, 500.0, 200.0), 30.0, (200.0, 500.0, 200.0), (500.0, 520.0, 700.0)))
    a, o = _pair()
    b, _ = _pair(terrain=True)
    seq = [a.edit(*e) for e in edits]
    for e in edits:
        o.edit(*e)
    bat, ms = b.edit_batch(edits)
    assert ms > 0
    for s1, s2 in zip(seq, bat):
        assert s1.applied == s2.applied
        assert list(s1.svo_items) == list(s2.svo_items)
    assert a.canonical() == b.canonical()
    assert b.canonical()[0] == o.canonical()[0]
    assert a.history() == b.history()
    a.close()
    b.close()
