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
        tp = dict(workloads.terrain_params(levels))
        tp["material"] = min(tp["material"], (1 << material_bits) - 1)
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
@pytest.mark.parametrize("bits,band,tile", [(8, 5, 0), (8, 5, 7), (4, 3, 8)])
def test_banded_terrain_and_copies_match_oracle(bits, band, tile):
    # Height-band materials, built in one edit (tile 0) or one edit per
    # column tile, then copy/paste edits with undo/redo, against the
    # oracle's factory rule and copy restatement.
    print(f"[case] banded terrain bits={bits} band={band} tile={tile}", flush=True)
    g = svdag_py.MyGPUHashDAG(L10, bits)
    o = oracle_py.OracleDag(L10, bits)
    tp = dict(workloads.terrain_params(L10))
    tp["material"] = 7 % (1 << bits)
    g.build_terrain(**tp, band_log2=band, band_seed=0xBA5E, tile_log2=tile)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"],
                    band, 0xBA5E)
    _same(g, o)
    hashes = [g.canonical()[0]]
    for src, dst, c, r in [((300.0, 500.0, 300.0), (600.0, 520.0, 650.0), (310.0, 505.0, 305.0), 40.0),
                           ((100.0, 470.0, 700.0), (140.0, 430.0, 690.0), (100.0, 480.0, 700.0), 31.0),
                           ((512.0, 400.0, 512.0), (256.0, 600.0, 768.0), (512.0, 420.0, 512.0), 64.0)]:
        st = g.edit(svdag_py.COPY, 0, c, r, src, dst)
        applied, counts = o.edit(svdag_py.COPY, 0, c, r, src, dst)
        assert st.applied == applied
        assert list(st.svo_items)[:L10 - 1] == counts[:L10 - 1]
        _same(g, o)
        hashes.append(g.canonical()[0])
    g.undo()
    g.undo()
    assert g.canonical()[0] == hashes[1]
    g.redo()
    assert g.canonical()[0] == hashes[2]
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [3, 5, 8])
def test_tiled_terrain_widens_tiles_to_the_widest_solid_node(tile):
    # A high base makes 512-wide nodes at y = 0 solid and fully filled.  A
    # build tiled narrower than that widens its tiles, so the DAG still equals
    # the factory rule's (one material: the band width is the whole volume).
    print(f"[case] tiled terrain tile={tile}", flush=True)
    g = svdag_py.MyGPUHashDAG(L10, 4)
    o = oracle_py.OracleDag(L10, 4)
    tp = dict(workloads.terrain_params(L10))
    tp.update(base=820.0, amplitude=90.0, material=5)
    g.build_terrain(**tp, band_log2=L10, band_seed=1, tile_log2=tile)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"], L10, 1)
    # The canonical dump lists levels bottom up: [level, count, (size, words)*]*.
    _, counts, dump = o.canonical(dump=True)
    p, full_at = 0, {}
    while p < len(dump) - 1:
        lvl, n = int(dump[p]), int(dump[p + 1])
        p += 2
        for _ in range(n):
            s = int(dump[p])
            if lvl < L10 - 2 and (int(dump[p + 1]) >> 8) & 1:
                full_at[lvl] = True
            p += 1 + s
    assert full_at.get(1) and not full_at.get(0)  # 512-wide nodes are full, the root is not
    _same(g, o)
    g.close()


ROTATED_COPIES = [
    # (src, dst, center, radius, quarter turns about +Y)
    ((300.0, 500.0, 300.0), (600.0, 520.0, 650.0), (300.0, 500.0, 300.0), 40.0, 1),
    ((512.0, 430.0, 512.0), (256.0, 560.0, 768.0), (512.0, 420.0, 512.0), 64.0, 2),
    ((100.5, 470.25, 700.0), (140.0, 430.5, 690.75), (101.0, 480.0, 699.0), 31.5, 3),
    ((620.0, 520.0, 640.0), (600.0, 520.0, 650.0), (610.0, 515.0, 655.0), 36.0, 1),  # overlapping regions
    ((40.0, 500.0, 980.0), (1010.0, 505.0, 20.0), (40.0, 500.0, 980.0), 30.0, 3),    # past the volume edge
]


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [8, 4])
def test_rotated_copies_match_oracle(bits):
    # Copy/paste with 90/180/270 degree turns about +Y on a height-band
    # terrain, one edit at a time and in a batch, with undo/redo: canonical
    # DAG and per-level edit-SVO sizes equal the oracle's after every copy.
    print(f"[case] rotated copies bits={bits}", flush=True)
    g = svdag_py.MyGPUHashDAG(L10, bits)
    o = oracle_py.OracleDag(L10, bits)
    tp = dict(workloads.terrain_params(L10))
    tp["material"] = 7 % (1 << bits)
    g.build_terrain(**tp, band_log2=4, band_seed=0xC0DE, tile_log2=0)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"],
                    4, 0xC0DE)
    _same(g, o)
    hashes = [g.canonical()[0]]
    for src, dst, c, r, k in ROTATED_COPIES:
        st = g.edit(svdag_py.COPY, 0, c, r, src, dst, k)
        applied, counts = o.edit(svdag_py.COPY, 0, c, r, src, dst, k)
        assert st.applied == applied
        assert list(st.svo_items)[:L10 - 1] == counts[:L10 - 1], (k, list(st.svo_items)[:L10 - 1], counts[:L10 - 1])
        _same(g, o)
        hashes.append(g.canonical()[0])
    assert len(set(hashes)) == len(hashes)
    for _ in range(3):
        g.undo()
    assert g.canonical()[0] == hashes[-4]
    g.redo()
    g.redo()
    assert g.canonical()[0] == hashes[-2]
    g.redo()
    # The same copies as one batch from the first copy's input agree too.
    o2 = oracle_py.OracleDag(L10, bits)
    o2.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"],
                     4, 0xC0DE)
    g2 = svdag_py.MyGPUHashDAG(L10, bits)
    g2.build_terrain(**tp, band_log2=4, band_seed=0xC0DE, tile_log2=0)
    batch = [(svdag_py.COPY, 0, c, r, src, dst, k) for src, dst, c, r, k in ROTATED_COPIES]
    batch.insert(2, (svdag_py.SPHERE_ADD, 3, (520.0, 440.0, 500.0), 20.0))
    stats, _ = g2.edit_batch(batch)
    for e, st in zip(batch, stats):
        applied, _ = o2.edit(e[0], e[1], e[2], e[3], *e[4:])
        assert st.applied == applied
    _same(g2, o2)
    g.close()
    g2.close()


@pytest.mark.gpu
def test_rotation_rejected_on_other_tools():
    g = svdag_py.MyGPUHashDAG(8, 4)
    with pytest.raises(svdag_py.SvdagError):
        g.edit(svdag_py.SPHERE_ADD, 1, (10.0, 10.0, 10.0), 4.0, rotation=1)
    with pytest.raises(svdag_py.SvdagError):
        g.edit(svdag_py.COPY, 0, (10.0, 10.0, 10.0), 4.0, (10.0, 10.0, 10.0), (40.0, 10.0, 10.0), 4)
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


def _mixed_stream(levels, seed):
    """Surface edits of all geometric kinds, clusters of colliding edits (the
    fused planner splits waves there), a copy, an edit outside the volume and
    an aligned fill that makes no leaf."""
    tp = workloads.terrain_params(levels)
    side = 1 << levels
    heights = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                        0, 0, side, side)

    def h(x, z):
        return float(heights[z, x])

    rng = np.random.default_rng(seed)
    edits = workloads.surface_edits(h, 40, seed=seed, levels=levels, rmin=2.0, rmax=side / 8, near=side / 16)
    for _ in range(8):
        x, z = (int(v) for v in rng.integers(0, side, 2))
        edits.insert(int(rng.integers(0, len(edits))),
                     (svdag_py.PAINT, int(rng.integers(1, 15)), (float(x), h(x, z), float(z)),
                      float(rng.uniform(3, side / 16))))
    c = edits[3][2]
    edits[5:5] = [(svdag_py.SPHERE_REMOVE, 0, c, 20.0), (svdag_py.BOX_ADD, 3, c, 9.0),
                  (svdag_py.SPHERE_ADD, 5, (c[0] + 7, c[1], c[2]), 12.0), (svdag_py.BOX_REMOVE, 0, c, 4.0)]
    src = (side * 0.3, h(int(side * 0.3), int(side * 0.3)), side * 0.3)
    dst = (side * 0.7, h(int(side * 0.7), int(side * 0.6)) + 5.0, side * 0.6)
    edits.insert(20, (svdag_py.COPY, 0, src, side / 32, src, dst))
    edits.append((svdag_py.SPHERE_ADD, 2, (-5000.0, 10.0, 10.0), 40.0))
    edits.append((svdag_py.BOX_ADD, 7, (side / 4, side / 4, side / 4), side / 4))
    return edits


@pytest.mark.gpu
@pytest.mark.parametrize("levels", [L10, 12])
def test_fused_waves_match_sequential_step_by_step(levels):
    edits = _mixed_stream(levels, seed=21 + levels)
    a, _ = _pair(levels=levels)
    b, o = _pair(levels=levels)
    a.set_batch_mode(0)
    b.set_batch_mode(1)
    sa, _ = a.edit_batch(edits[:30])
    sb, _ = b.edit_batch(edits[:30])
    sa2, _ = a.edit_batch(edits[30:])
    sb2, _ = b.edit_batch(edits[30:])
    for k, (s1, s2) in enumerate(zip(sa + sa2, sb + sb2)):
        assert s1.applied == s2.applied, k
        assert list(s1.svo_items) == list(s2.svo_items), k
    for e in edits:
        o.edit(*e)
    _same(b, o)
    assert a.history() == b.history()
    for _ in range(a.history()[0]):
        assert a.canonical() == b.canonical()
        a.undo()
        b.undo()
    assert a.canonical() == b.canonical()
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("gap", [40.0, 70.0, 130.0, 260.0])
def test_paint_waves_match_sequential_when_earlier_edits_change_emptiness(gap):
    # Sphere adds in the empty sky above the terrain, each followed by a paint
    # `gap` voxels away: the paint's cells are its own, but its spine nodes
    # are the ones the add made non-empty (and a remove then a paint on the
    # surface, the other way round).  The fused form must give the sequential
    # form's SVO sizes and applied flags edit by edit, and the oracle's DAG.
    levels = L10
    tp = workloads.terrain_params(levels)
    side = 1 << levels
    heights = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                        0, 0, side, side)
    edits = []
    for j, x in enumerate((150.0, 450.0, 750.0)):
        edits.append((svdag_py.SPHERE_ADD, 3 + j, (x, 940.0, 300.0), 18.0))
        edits.append((svdag_py.PAINT, 7 + j, (x, 940.0, 300.0 + gap), 18.0))
        edits.append((svdag_py.PAINT, 9, (x, 940.0, 300.0 - gap), 25.0))
    for x in (200.0, 600.0):
        y = float(heights[600, int(x)])
        edits.append((svdag_py.SPHERE_REMOVE, 0, (x, y, 600.0), 30.0))
        edits.append((svdag_py.PAINT, 12, (x + gap, y, 600.0), 30.0))
    a, _ = _pair()
    b, o = _pair()
    a.set_batch_mode(0)
    b.set_batch_mode(1)
    sa, _ = a.edit_batch(edits)
    sb, _ = b.edit_batch(edits)
    for k, (s1, s2) in enumerate(zip(sa, sb)):
        assert s1.applied == s2.applied, k
        assert list(s1.svo_items) == list(s2.svo_items), k
    for e in edits:
        o.edit(*e)
    _same(b, o)
    assert a.canonical() == b.canonical()
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [16, 64])
def test_parallel_spine_waves_match_sequential(batch):
    # Sphere and box streams only (no paint): their waves take the parallel
    # spine.  Every edit's applied flag, SVO sizes and resulting DAG (walked
    # back through undo) equal the sequential form; colliding clusters, an
    # off-volume edit and a no-leaf fill are mixed in.
    levels = L10
    tp = workloads.terrain_params(levels)
    side = 1 << levels
    heights = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                        0, 0, side, side)
    edits = workloads.surface_edits(lambda x, z: float(heights[z, x]), 72, seed=31 + batch, levels=levels,
                                    rmin=2.0, rmax=side / 32, near=side / 32)
    c = edits[3][2]
    edits[5:5] = [(svdag_py.SPHERE_REMOVE, 0, c, 20.0), (svdag_py.BOX_ADD, 3, c, 9.0),
                  (svdag_py.SPHERE_ADD, 5, (c[0] + 7, c[1], c[2]), 12.0)]
    edits.insert(40, (svdag_py.SPHERE_ADD, 2, (-5000.0, 10.0, 10.0), 40.0))
    edits.insert(41, (svdag_py.BOX_ADD, 7, (side / 4, side / 4, side / 4), side / 4))
    a, _ = _pair(levels=levels)
    b, o = _pair(levels=levels)
    a.set_batch_mode(0)
    b.set_batch_mode(1)
    sa, sb = [], []
    for i in range(0, len(edits), batch):
        sa += a.edit_batch(edits[i:i + batch])[0]
        sb += b.edit_batch(edits[i:i + batch])[0]
    for k, (s1, s2) in enumerate(zip(sa, sb)):
        assert s1.applied == s2.applied, k
        assert list(s1.svo_items) == list(s2.svo_items), k
    for e in edits:
        o.edit(*e)
    _same(b, o)
    assert a.history() == b.history()
    for _ in range(a.history()[0]):
        assert a.canonical() == b.canonical()
        a.undo()
        b.undo()
    assert a.canonical() == b.canonical()
    a.close()
    b.close()


def _frame_camera(levels):
    cam = workloads.config1_camera(levels)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    return (cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"])


@pytest.mark.gpu
def test_serving_results_equal_alone_and_in_the_mix():
    # Requests arrive faster than batches finish, so batches mix many of
    # them; every request's result (applied flag and the DAG state after it)
    # equals applying the requests one by one on their own.
    edits = _mixed_stream(L10, seed=33)
    rng = np.random.default_rng(3)
    arrivals = np.cumsum(rng.exponential(1 / 3000.0, len(edits)))
    a, o = _pair()
    b, _ = _pair()
    cam = _frame_camera(L10)
    info, st = a.serve_replay(edits, arrivals, max_batch=16, frame=(256, 256), cameras=[cam], frame_period_us=1000)
    assert (info["status"] == 0).all()
    assert st.requests == len(edits) and st.batches >= 2 and st.max_batch_seen > 1
    assert st.frames > 0
    assert (info["done_s"] >= info["submit_s"]).all()
    alone = [b.edit(*e).applied for e in edits]
    assert list(info["applied"]) == alone
    for e in edits:
        o.edit(*e)
    _same(a, o)
    assert a.history() == b.history()
    roots = [a.root]
    for _ in range(a.history()[0]):
        assert a.canonical() == b.canonical()
        a.undo()
        b.undo()
        roots.append(a.root)
    assert a.canonical() == b.canonical()
    # The last frame traced a published state: tracing that state again
    # gives the same hit buffer.
    hits, froot = a.last_frame(256, 256)
    assert froot in roots
    for _ in range(len(roots)):
        if a.root == froot:
            break
        a.redo()
    assert a.root == froot
    again, _ = a.trace(*cam, 256, 256)
    assert np.array_equal(hits.view(np.uint32), again.view(np.uint32))
    a.close()
    b.close()


@pytest.mark.gpu
def test_serving_admission_collects_and_grows_under_a_small_arena():
    # Arenas far too small for the stream: admission has to collect garbage
    # and grow the arenas, and the result still equals the oracle's.
    a, o = _pair(expected_items=20000, arena_slabs=64)
    edits = _mixed_stream(L10, seed=41)
    arrivals = np.arange(len(edits)) * 2e-4
    info, st = a.serve_replay(edits, arrivals, max_batch=8, allow_gc=True, reserve_fraction=0.05)
    assert (info["status"] == 0).all()
    assert st.gcs + st.grows > 0
    for e in edits:
        o.edit(*e)
    _same(a, o)
    m = a.memory_usage()
    assert m.slabs_used + m.slabs_free == m.slab_capacity
    a.close()


@pytest.mark.gpu
def test_autotune_keeps_the_dag_and_later_edits_exact():
    a, o = _pair()
    edits = _mixed_stream(L10, seed=55)
    before = a.canonical()
    g, ms = a.autotune(edits[:16])
    assert g in (4, 8, 16, 32) and all(v > 0 for v in ms.values())
    assert a.canonical() == before and a.history() == (0, 0)
    a.edit_batch(edits)
    for e in edits:
        o.edit(*e)
    _same(a, o)
    a.close()


@pytest.mark.gpu
def test_autotune_all_picks_both_knobs_and_keeps_the_dag():
    # Lanes per item, then the per-level grid cap: both picks come from the
    # candidate sets, every candidate was timed, the DAG is unchanged, and
    # later fused and single edits under the picks equal the oracle.
    a, o = _pair()
    edits = _mixed_stream(L10, seed=57)
    before = a.canonical()
    rep = a.autotune_all(edits[:16])
    assert rep["lanes_per_item"] in (4, 8, 16, 32)
    assert rep["max_blocks"] in (1024, 2048, 4096, 8192, 16384)
    assert all(v > 0 for v in rep["ms_by_lanes"].values())
    assert all(v > 0 for v in rep["ms_by_max_blocks"].values())
    assert a.canonical() == before and a.history() == (0, 0)
    a.edit_batch(edits[:40])
    for e in edits[40:]:
        a.edit(*e)
    for e in edits:
        o.edit(*e)
    _same(a, o)
    a.close()


@pytest.mark.gpu
def test_set_load_factor_rebuilds_the_tables_and_keeps_the_dag():
    # MyGPUHashDAG::setLoadFactor: at every load factor the rebuilt tables
    # hold the same canonical DAG, lower load factors use more head lines,
    # and edits, undo / redo and collections after a rebuild equal the oracle.
    a, o = _pair()
    edits = _mixed_stream(L10, seed=71)
    for e in edits[:20]:
        a.edit(*e)
        o.edit(*e)
    before = a.canonical()
    used = {}
    for lf in (30, 95, 60):
        a.set_load_factor(lf)
        assert a.canonical() == before and a.history() == (0, 0)
        used[lf] = a.memory_usage().slabs_used
    assert used[30] > used[60] > used[95]
    a.garbage_collect()
    assert a.canonical() == before
    a.edit_batch(edits[20:36])
    for e in edits[36:44]:
        a.edit(*e)
    # undo, undo, redo: the DAG after edits[:43].
    a.undo()
    a.undo()
    a.redo()
    for e in edits[20:43]:
        o.edit(*e)
    _same(a, o)
    assert a.garbage_collect().items_live > 0
    _same(a, o)
    a.close()


@pytest.mark.gpu
def test_autotune_load_factor_picks_a_candidate_and_keeps_the_dag():
    a, o = _pair()
    edits = _mixed_stream(L10, seed=73)
    before = a.canonical()
    rep = a.autotune_load_factor(edits[:16], (40, 75, 95))
    assert rep["load_factor"] in (40, 75, 95)
    assert all(v > 0 for v in rep["ms_by_load_factor"].values())
    assert a.canonical() == before
    a.edit_batch(edits[:32])
    for e in edits[:32]:
        o.edit(*e)
    _same(a, o)
    a.close()


@pytest.mark.gpu
def test_graph_waves_replay_exactly_across_batches():
    # Fused waves replayed from captured HIP graphs: several batches reuse one
    # graph per split level, with eager work (single edits, a collection that
    # clears the history) between the replays; every batch's states match the
    # sequential form edit by edit.
    edits = _mixed_stream(L10, seed=61)
    a, _ = _pair()
    b, o = _pair()
    a.set_batch_mode(0)
    b.set_batch_mode(2)
    chunks = [edits[0:12], edits[12:24], edits[24:36], edits[36:]]
    for i, chunk in enumerate(chunks):
        sa, _ = a.edit_batch(chunk)
        sb, _ = b.edit_batch(chunk)
        for s1, s2 in zip(sa, sb):
            assert s1.applied == s2.applied
            assert list(s1.svo_items) == list(s2.svo_items)
        assert b.history() == a.history()
        for _ in range(min(len(chunk), a.history()[0])):
            assert a.canonical() == b.canonical()
            a.undo()
            b.undo()
        for _ in range(a.history()[1]):
            a.redo()
            b.redo()
        assert a.canonical() == b.canonical()
        # eager work between replays
        extra = (svdag_py.SPHERE_ADD, 3, (300.0 + 40 * i, 560.0, 420.0), 25.0)
        a.edit(*extra)
        b.edit(*extra)
        if i == 1:
            a.garbage_collect()
            b.garbage_collect()
    for i, chunk in enumerate(chunks):
        for e in chunk:
            o.edit(*e)
        o.edit(svdag_py.SPHERE_ADD, 3, (300.0 + 40 * i, 560.0, 420.0), 25.0)
    _same(b, o)
    assert a.canonical() == b.canonical()
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("material_bits", [0, 8])
def test_other_material_widths_match_oracle(material_bits):
    # 8 material bits: leaves of up to 18 words (the wide needle kernels);
    # 0 bits: geometry only, one leaf size.  Edits of every kind, fused
    # batches, undo, and a traced frame.
    levels = L10
    g, o = _pair(material_bits=material_bits)
    nm = 1 << material_bits
    edits = _mixed_stream(levels, seed=90 + material_bits)
    edits = [(k, (m * 37) % nm if k != svdag_py.SPHERE_REMOVE else 0, *rest) for k, m, *rest in edits]
    g.edit_batch(edits[:30])
    for e in edits[30:]:
        g.edit(*e)
    for e in edits:
        o.edit(*e)
    _same(g, o)
    g.undo()
    g.undo()
    g.redo()
    cam = workloads.config1_camera(levels)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    gh, _ = g.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 128, 128)
    g.redo()
    _same(g, o)
    oh = o.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 128, 128)
    gh2, _ = g.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 128, 128)
    assert np.array_equal(gh2.view(np.uint32), oh.view(np.uint32))
    g.close()


@pytest.mark.gpu
def test_marking_collection_retries_after_frontier_overflow(monkeypatch):
    # With the first attempt's frontier capped at 4096 entries the mark
    # overflows at several levels and reruns with doubled frontiers; a level
    # after an overflow must expand nothing (its input has an unwritten
    # hole).  The result must equal an unforced collection's: a marking one
    # right after frees nothing, and the DAG stays equal to the oracle's.
    levels = L10
    edits = _mixed_stream(levels, seed=5)
    g, o = _pair(levels=levels)
    g.edit_batch(edits)
    for e in edits:
        o.edit(*e)
    monkeypatch.setenv("SVDAG_GC_FRONTIER_ENTRIES", "4096")
    forced = g.garbage_collect(mark=True)
    monkeypatch.delenv("SVDAG_GC_FRONTIER_ENTRIES")
    assert forced.items_freed > 0
    assert forced.items_live == _stored_items(g)
    check = g.garbage_collect(mark=True)
    assert check.items_freed == 0 and check.items_live == forced.items_live
    _same(g, o)
    # Some level's frontier is larger than the cap, so the retries ran.
    rec = g.export_records()
    per_level, p = {}, 4
    for _ in range(int(rec[3])):
        level, handle, size = int(rec[p]), int(rec[p + 1]), int(rec[p + 2])
        per_level.setdefault(level, set()).add(handle)
        p += 3 + size
    assert max(len(v) for lv, v in per_level.items() if lv < g.levels - 2) > 4096
    g.close()


@pytest.mark.gpu
def test_counting_collection_marks_after_growth():
    # A count that reached its ceiling never drops, so only a marking
    # collection frees such an item.  Once the stored items exceed 1.5 times
    # the live count of the last marking collection, svdag_gc marks
    # (stats.marked), and the next one counts again.  12 sphere adds of new
    # materials grow the depth-10 terrain's distinct items 1.79x (oracle).
    g, o = _pair()
    base = g.garbage_collect()
    tp = workloads.terrain_params(L10)
    heights = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                        0, 0, 1024, 1024)
    rng = np.random.default_rng(3)
    for i in range(12):
        x, z = rng.integers(64, 960, 2)
        e = (svdag_py.SPHERE_ADD, 1 + i % 14, (float(x), float(heights[z, x]), float(z)), float(rng.uniform(50, 90)))
        g.edit(*e)
        o.edit(*e)
    grown = g.garbage_collect()
    assert grown.marked == 0 and grown.items_live > 1.5 * base.items_live
    again = g.garbage_collect()
    assert again.marked == 1 and again.items_live <= grown.items_live
    assert again.items_live == _stored_items(g)
    after = g.garbage_collect()
    assert after.marked == 0 and after.items_freed == 0 and after.items_live == again.items_live
    _same(g, o)
    g.close()


@pytest.mark.gpu
def test_reference_count_collection_frees_what_marking_frees():
    # svdag_gc releases the roots held since the last collection and frees
    # what drops to zero; svdag_gc_mark marks from the roots, sweeps and
    # recounts.  After every counting collection a marking one must free
    # nothing and agree on the live count, which must equal the distinct
    # stored items reachable from the roots; the DAG stays equal to the
    # oracle's through edits in both batch modes, undo / redo, a copy and
    # a paint, across several collections of each kind.
    levels = L10
    edits = _mixed_stream(levels, seed=77)
    g, o = _pair(levels=levels)
    first = g.garbage_collect()
    check = g.garbage_collect(mark=True)
    assert check.items_freed == 0 and check.items_live == first.items_live == _stored_items(g)
    chunks = [edits[:12], edits[12:30], edits[30:]]
    for step, chunk in enumerate(chunks):
        g.set_batch_mode(step % 2)
        g.edit_batch(chunk)
        for e in chunk:
            o.edit(*e)
        if step == 1:
            # Roots visited by undo / redo are released too.
            g.undo()
            g.undo()
            g.redo()
            g.redo()
        st = g.garbage_collect()
        assert st.items_freed > 0 and st.marked == 0
        assert st.items_live == _stored_items(g)
        _same(g, o)
        check = g.garbage_collect(mark=True)
        assert check.items_freed == 0 and check.marked == 1, (step, check.items_freed)
        assert check.items_live == st.items_live, step
    # An abandoned undo branch: its roots die at the next collection.
    g.edit(svdag_py.SPHERE_ADD, 3, (300.0, 520.0, 300.0), 40.0)
    g.undo()
    g.edit(svdag_py.SPHERE_REMOVE, 0, (700.0, 500.0, 700.0), 30.0)
    o.edit(svdag_py.SPHERE_REMOVE, 0, (700.0, 500.0, 700.0), 30.0)
    st = g.garbage_collect()
    assert st.items_live == _stored_items(g)
    check = g.garbage_collect(mark=True)
    assert check.items_freed == 0 and check.items_live == st.items_live
    _same(g, o)
    g.close()


def _fusable_stream(levels, n, seed):
    """Sphere / box / paint edits small enough for fused waves (split level
    at least 2), near the terrain surface."""
    tp = workloads.terrain_params(levels)
    side = 1 << levels
    heights = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                        0, 0, side, side)

    def h(x, z):
        return float(heights[z, x])

    edits = workloads.surface_edits(h, n, seed=seed, levels=levels, rmin=2.0, rmax=side / 32, near=side / 16)
    rng = np.random.default_rng(seed)
    for _ in range(n // 8):
        x, z = (int(v) for v in rng.integers(0, side, 2))
        edits.insert(int(rng.integers(0, len(edits))),
                     (svdag_py.PAINT, int(rng.integers(1, 15)), (float(x), h(x, z), float(z)),
                      float(rng.uniform(3, side / 32))))
    return edits


@pytest.mark.gpu
@pytest.mark.parametrize("collect,mode", [(True, 1), (False, 1), (True, 2)])
def test_edit_stream_matches_batches_and_oracle(collect, mode):
    # svdag_edit_stream enqueues every batch (and its collection) before
    # waiting; in fused mode each collection runs beside the next batch's
    # SVO build.  The DAG, the per-edit results and what is stored afterwards
    # equal svdag_edit_batch per batch with svdag_gc after each, and the
    # oracle's DAG.  The last batch is short (81 edits in batches of 8).
    levels = L10
    edits = _fusable_stream(levels, 72, seed=23)
    g, o = _pair(levels=levels)
    w, _ = _pair(levels=levels)
    g.set_batch_mode(mode)
    g.garbage_collect()
    w.garbage_collect()
    st, ms = g.edit_stream(edits, batch=8, collect=collect)
    sw = []
    for b in range(0, len(edits), 8):
        sw += w.edit_batch(edits[b:b + 8])[0]
        if collect:
            w.garbage_collect()
    for e in edits:
        o.edit(*e)
    _same(g, o)
    assert ms > 0
    assert [s.applied for s in st] == [s.applied for s in sw]
    assert [list(s.svo_items) for s in st] == [list(s.svo_items) for s in sw]
    if collect:
        assert g.history() == (0, 0)
        live, lw = g.garbage_collect(), w.garbage_collect()
        print(f"[case] live after stream {live.items_live}, reachable {_stored_items(g)}; "
              f"per batch {lw.items_live}, reachable {_stored_items(w)}", flush=True)
        # The stream's collections released every root it replaced, the
        # last (short) batch's included: nothing is left to free, and each
        # context's stored items are exactly what its roots reach.
        assert live.items_freed == 0
        assert live.marked == 0 and live.items_live == _stored_items(g)
        assert lw.items_live == _stored_items(w)
        check = g.garbage_collect(mark=True)
        assert check.items_freed == 0 and check.items_live == live.items_live
    else:
        assert g.history() == w.history()
    g.close()
    w.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [1, 5, 64])
def test_edit_stream_batch_sizes(batch):
    # Batches of one edit, a short last batch, and one batch larger than the
    # stream: the DAG equals the oracle's, every released root was collected
    # (a collection afterwards frees nothing), and the stored items are what
    # the root reaches.
    levels = L10
    edits = _fusable_stream(levels, 18, seed=61)
    g, o = _pair(levels=levels)
    g.garbage_collect()
    st, _ = g.edit_stream(edits, batch=batch, collect=True)
    assert len(st) == len(edits)
    for e in edits:
        o.edit(*e)
    _same(g, o)
    after = g.garbage_collect()
    assert after.items_freed == 0 and after.items_live == _stored_items(g)
    assert svdag_py.lib().svdag_edit_stream(g.h, None, 0, batch, 1, None, None) == 0
    g.close()


@pytest.mark.gpu
def test_add_frame_follows_the_reference_undo_stack():
    # MyGPUHashDAGUndoRedo::add_frame: the frame goes on the undo stack unless
    # it is the last one there, and redo is cleared; undo then returns to it.
    g, _ = _pair()
    r0 = g.root
    g.edit(svdag_py.SPHERE_ADD, 3, (300.0, 520.0, 300.0), 30.0)
    ra = g.root
    assert g.history() == (1, 0)
    g.add_frame(ra)
    g.add_frame(ra)  # the last frame already: not pushed again
    assert g.history() == (2, 0)
    g.undo()
    assert g.root == ra and g.history() == (1, 1)
    g.undo()
    assert g.root == r0 and g.history() == (0, 2)
    g.add_frame(r0)  # clears redo
    assert g.history() == (1, 0)
    with pytest.raises(svdag_py.SvdagError):
        g.add_frame(0x1234567)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2])
def test_edit_stream_recreates_what_the_previous_collection_frees(mode):
    # Batches that add, remove and add again the same spheres: each
    # collection frees the nodes the next batch then builds again, while that
    # batch's SVO build runs beside the collection and its merge waits for
    # it.  The DAG equals the oracle's after every step, and a collection
    # after the stream frees nothing.
    levels = L10
    side = 1 << levels
    rng = np.random.default_rng(97)
    centers = [tuple(float(v) for v in rng.uniform(100, side - 100, 3)) for _ in range(6)]
    add = [(svdag_py.SPHERE_ADD, 3 + k, c, 20.0 + 4 * k) for k, c in enumerate(centers)]
    rem = [(svdag_py.SPHERE_REMOVE, 0, c, 20.0 + 4 * k) for k, c in enumerate(centers)]
    edits = add + rem + add + rem + add
    g, o = _pair(levels=levels)
    g.set_batch_mode(mode)
    g.garbage_collect()
    g.edit_stream(edits, batch=6, collect=True)
    for e in edits:
        o.edit(*e)
    _same(g, o)
    st = g.garbage_collect()
    assert st.items_freed == 0 and st.items_live == _stored_items(g)
    check = g.garbage_collect(mark=True)
    assert check.items_freed == 0 and check.items_live == st.items_live
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2])
def test_edit_stream_recovers_when_an_arena_runs_out(mode):
    # Arenas far too small: an edit of the pipelined stream runs out on the
    # device, the stream's later collections release nothing, and the rest
    # runs batch by batch after the arenas grow.  The DAG equals the
    # oracle's, and the next collection marks and frees what the abandoned
    # part inserted.
    levels = L10
    edits = _fusable_stream(levels, 96, seed=29)
    g, o = _pair(levels=levels, expected_items=20000, arena_slabs=64)
    g.set_batch_mode(mode)
    g.garbage_collect()
    cap0 = g.memory_usage().slab_capacity
    g.edit_stream(edits, batch=8, collect=True)
    assert g.memory_usage().slab_capacity > cap0
    for e in edits:
        o.edit(*e)
    _same(g, o)
    st = g.garbage_collect()
    assert st.items_live == _stored_items(g)
    check = g.garbage_collect(mark=True)
    assert check.items_freed == 0 and check.items_live == st.items_live
    g.close()

@pytest.mark.gpu
def test_deferred_collections_free_what_waited_ones_free():
    # svdag_gc without stats only enqueues the counting collection.  A
    # stream of batches, each followed by a deferred collection, must leave
    # the same DAG as the oracle and the same stored items as the same
    # stream with waited collections; a marking collection afterwards frees
    # nothing.  Queries between deferred collections see the edited DAG.
    levels = L10
    edits = _mixed_stream(levels, seed=91)
    g, o = _pair(levels=levels)
    w, _ = _pair(levels=levels)
    g.garbage_collect()
    w.garbage_collect()
    for b in range(0, len(edits), 8):
        chunk = edits[b:b + 8]
        g.edit_batch(chunk)
        w.edit_batch(chunk)
        for e in chunk:
            o.edit(*e)
        assert g.garbage_collect(wait=False) is None
        w.garbage_collect()
        pts = np.random.default_rng(b).integers(0, 1 << levels, (2000, 3), dtype=np.uint32)
        assert np.array_equal(g.query(pts), o.query(pts)), b
    _same(g, o)
    st = g.garbage_collect()
    sw = w.garbage_collect()
    assert st.items_live == sw.items_live == _stored_items(g) == _stored_items(w)
    check = g.garbage_collect(mark=True)
    assert check.items_freed == 0 and check.items_live == st.items_live
    g.close()
    w.close()


@pytest.mark.gpu
def test_config3_stream_depth13_fused_and_graph_match_oracle():
    # The benchmark's batched-editing generator (40% sphere add, 40% sphere
    # remove, 20% box fill, r log-uniform in [8, 512], centres within +-1024
    # of the surface) at depth 13: 64 edits in fused batches of 16 with a
    # collection after every batch, in eager fused mode and in graph mode,
    # each equal to the oracle's sequential edits.
    levels = 13
    g, o = _pair(levels=levels, expected_items=4_000_000, workspace_items=1 << 24)
    side = 1 << levels
    tp = workloads.terrain_params(levels)
    hts = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                    0, 0, side, side)
    edits = workloads.config3_edits(lambda x, z: int(hts[z, x]), levels, n=64, seed=6)
    for e in edits:
        o.edit(*e)
    oh, oc = o.canonical()
    for mode in (1, 2):
        if mode == 2:
            g.close()
            g = svdag_py.MyGPUHashDAG(levels, 4, expected_items=4_000_000, workspace_items=1 << 24)
            g.build_terrain(**tp)
        g.set_batch_mode(mode)
        g.autotune(edits[:16])
        g.garbage_collect()
        for b in range(0, len(edits), 16):
            st, _ = g.edit_batch(edits[b:b + 16])
            assert len(st) == 16
            gc = g.garbage_collect()
            assert gc.marked == 0, (mode, b)
        gh, gcnt = g.canonical()
        assert gcnt == oc and gh == oh, mode
    g.close()
