"""GPU slab hash tables (lane and wave64-group variants) against the oracle's
restatement of CompactAccelerationHashTable: set semantics, bit-exact words."""
import numpy as np
import pytest

import oracle_py
import svdag_py
import workloads

VARIANTS = [0, 4, 8, 16, 32]


def _build(items, fill=0.75, reserve_scale=1):
    sizes = list(range(2, 10))
    b = workloads.buckets_for(items, fill)
    hs = svdag_py.HashSet(2, 9, [b.get(s, 1) for s in sizes],
                          [reserve_scale * len(items.get(s, [])) for s in sizes])
    return hs, b


def _oracle_found(items, finds, buckets):
    out = {}
    for s, a in items.items():
        t = oracle_py.OracleTable(s, buckets[s])
        t.add(a)
        out[s] = t.find(finds[s]) != 0xFFFFFFFF
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("variant", VARIANTS)
def test_bulk_add_find_parity(variant):
    items = workloads.slab_hash_items(1 << 17)
    finds = workloads.slab_hash_finds(items, 1 << 17)
    hs, buckets = _build(items)
    groups = sorted(items.items())
    words, segs, n = svdag_py.make_segments(groups)
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    hs.add_dev(wdev, segs, odev, variant)
    ptrs = odev.to_array(np.uint32, n)
    assert np.all(ptrs != svdag_py.NOT_FOUND)
    # Every insert decodes back to its own words.
    for i, (s, a) in enumerate(groups):
        seg = segs[i]
        got = hs.read(s, ptrs[seg.out_offset:seg.out_offset + seg.count])
        assert np.array_equal(got, a), s
    # Finds: present exactly when the oracle finds the item, and the found
    # pointer decodes to the needle.
    fgroups = [(s, finds[s]) for s, _ in groups]
    fwords, fsegs, fn = svdag_py.make_segments(fgroups)
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    rdev = svdag_py.DeviceBuffer(4 * fn)
    hs.find_dev(fdev, fsegs, rdev, variant)
    res = rdev.to_array(np.uint32, fn)
    expect = _oracle_found(items, finds, buckets)
    for i, (s, a) in enumerate(fgroups):
        seg = fsegs[i]
        r = res[seg.out_offset:seg.out_offset + seg.count]
        hit = r != svdag_py.NOT_FOUND
        assert np.array_equal(hit, expect[s]), s
        assert np.array_equal(hs.read(s, r[hit]), a[hit]), s
    hs.close()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [1, 2, 3, 204, 408, 1004, 1208])
def test_find_only_variants_match_oracle_at_2_20(variant):
    # The find-only variants (the benchmark's timed one among them) against
    # the oracle on 2^20 items and 2^20 finds at the benchmark's 75% fill:
    # hit exactly where the oracle hits, and every hit decodes to its needle.
    items = workloads.slab_hash_items(1 << 20)
    finds = workloads.slab_hash_finds(items, 1 << 20, hit_rate=0.5)
    hs, buckets = _build(items)
    groups = sorted(items.items())
    words, segs, n = svdag_py.make_segments(groups)
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    hs.add_dev(wdev, segs, odev, 16)
    fgroups = [(s, finds[s]) for s, _ in groups]
    fwords, fsegs, fn = svdag_py.make_segments(fgroups)
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    rdev = svdag_py.DeviceBuffer(4 * fn)
    hs.find_dev(fdev, fsegs, rdev, variant)
    res = rdev.to_array(np.uint32, fn)
    expect = _oracle_found(items, finds, buckets)
    hits = 0
    for i, (s, a) in enumerate(fgroups):
        seg = fsegs[i]
        r = res[seg.out_offset:seg.out_offset + seg.count]
        hit = r != svdag_py.NOT_FOUND
        assert np.array_equal(hit, expect[s]), s
        assert np.array_equal(hs.read(s, r[hit]), a[hit]), s
        hits += int(hit.sum())
    assert abs(hits / fn - 0.5) < 0.01
    hs.close()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 16])
def test_long_chains_and_duplicates(variant):
    # Few buckets force multi-slab chains; a small alphabet forces duplicates.
    rng = np.random.default_rng(7)
    items = {3: rng.integers(0, 40, size=(20000, 3)).astype(np.uint32)}
    items[3][:, 0] = 3  # valid two-child header
    hs = svdag_py.HashSet(3, 3, [5], [len(items[3]) * 8])
    words, segs, n = svdag_py.make_segments([(3, items[3])])
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    hs.add_dev(wdev, segs, odev, variant)
    ptrs = odev.to_array(np.uint32, n)
    assert len(set(ptrs.tolist())) == n  # multiset insert: one slot per item
    assert np.array_equal(hs.read(3, ptrs), items[3])
    probes = rng.integers(0, 45, size=(5000, 3)).astype(np.uint32)
    probes[:, 0] = 3
    pw, ps, pn = svdag_py.make_segments([(3, probes)])
    pdev = svdag_py.DeviceBuffer.from_array(pw)
    rdev = svdag_py.DeviceBuffer(4 * pn)
    hs.find_dev(pdev, ps, rdev, variant)
    r = rdev.to_array(np.uint32, pn)
    present = {tuple(x) for x in items[3].tolist()}
    assert [bool(x != svdag_py.NOT_FOUND) for x in r] == [tuple(p) in present for p in probes.tolist()]
    hs.close()


@pytest.mark.gpu
def test_queued_finds_equal_waited_finds():
    # Untimed finds are only queued on the set's stream: several back to back
    # into different result buffers, then the region timer's wait, give what
    # a waited find gives, and the timer reports a positive device time.
    items = workloads.slab_hash_items(1 << 18)
    finds = workloads.slab_hash_finds(items, 1 << 18, hit_rate=0.5)
    hs, _ = _build(items)
    groups = sorted(items.items())
    words, segs, n = svdag_py.make_segments(groups)
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    hs.add_dev(wdev, segs, odev, 16)
    fwords, fsegs, fn = svdag_py.make_segments([(s, finds[s]) for s, _ in groups])
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    ref = svdag_py.DeviceBuffer(4 * fn)
    hs.find_dev(fdev, fsegs, ref, 1)
    outs = [svdag_py.DeviceBuffer(4 * fn) for _ in range(3)]
    hs.timer(False)
    for o in outs:
        assert hs.find_dev(fdev, fsegs, o, 1, timed=False) is None
    assert hs.timer(True) > 0
    expect = ref.to_array(np.uint32, fn)
    for o in outs:
        assert np.array_equal(o.to_array(np.uint32, fn), expect)
    hs.close()


@pytest.mark.gpu
def test_empty_batch_and_oom_is_reported():
    hs = svdag_py.HashSet(2, 2, [1], [0])
    usage = hs.usage(2)
    assert usage[0] == 1 and usage[2] == 1
    # 1 head line + a reserve of 64 overflow lines of 13 two-word items: 13 * 65 fit, more do not.
    cap_items = workloads.line_slots(2) * usage[1]
    items = np.stack([np.full(cap_items + 500, 1, np.uint32),
                      np.arange(cap_items + 500, dtype=np.uint32)], axis=1)
    words, segs, n = svdag_py.make_segments([(2, items)])
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    with pytest.raises(svdag_py.SvdagError) as e:
        hs.add_dev(wdev, segs, odev, 16)
    assert e.value.code == -2
    hs.close()


FIND_VARIANTS = [0, 1, 2, 3, 4, 8, 16, 32, 204, 404, 208, 408, 1004, 1008, 1204, 1208]


@pytest.mark.gpu
@pytest.mark.parametrize("buckets_per_size", [None, 3])
def test_find_variants_return_identical_pointers(buckets_per_size):
    # All find variants (lane, group, multi-needle, XCD-binned) run the same
    # probe over the same table: the result arrays are identical, including
    # multi-slab chains (3 buckets per size).
    items = workloads.slab_hash_items(1 << 16)
    finds = workloads.slab_hash_finds(items, 1 << 16)
    if buckets_per_size is None:
        hs, _ = _build(items)
    else:
        sizes = list(range(2, 10))
        hs = svdag_py.HashSet(2, 9, [buckets_per_size] * len(sizes),
                              [len(items.get(s, [])) * 2 for s in sizes])
    groups = sorted(items.items())
    words, segs, n = svdag_py.make_segments(groups)
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    hs.add_dev(wdev, segs, odev, 16)
    fwords, fsegs, fn = svdag_py.make_segments([(s, finds[s]) for s, _ in groups])
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    rdev = svdag_py.DeviceBuffer(4 * fn)
    ref = None
    for v in FIND_VARIANTS:
        hs.find_dev(fdev, fsegs, rdev, v)
        r = rdev.to_array(np.uint32, fn)
        if ref is None:
            ref = r
            assert 0.4 < float(np.mean(r != svdag_py.NOT_FOUND)) < 0.6
        assert np.array_equal(r, ref), v
    hs.close()


@pytest.mark.gpu
def test_memory_probes_run():
    # Stream and random-line probes (bench.py's practical ceilings) return a
    # positive kernel time; bad arguments are rejected.
    assert svdag_py.hbm_probe(svdag_py.PROBE_STREAM, 64 << 20) > 0.0
    assert svdag_py.hbm_probe(svdag_py.PROBE_RANDOM_LINES, 64 << 20, 1 << 20) > 0.0
    with pytest.raises(svdag_py.SvdagError):
        svdag_py.hbm_probe(7, 64 << 20)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [1, 2, 3])
def test_lane_segment_variants_ragged_and_empty_segments(variant):
    # Blocks are assigned to segments on the host: an empty segment, a
    # one-item segment and a ragged one must still give the per-lane result.
    rng = np.random.default_rng(11)
    items = {3: rng.integers(0, 1 << 32, size=(1000, 3), dtype=np.uint64).astype(np.uint32),
             5: rng.integers(0, 1 << 32, size=(777, 5), dtype=np.uint64).astype(np.uint32)}
    items[3][:, 0] = (items[3][:, 0] & ~np.uint32(0xFF)) | np.uint32(3)
    items[5][:, 0] = (items[5][:, 0] & ~np.uint32(0xFF)) | np.uint32(15)
    hs = svdag_py.HashSet(2, 9, [7] * 8, [2000] * 8)
    words, segs, n = svdag_py.make_segments(sorted(items.items()))
    wdev = svdag_py.DeviceBuffer.from_array(words)
    odev = svdag_py.DeviceBuffer(4 * n)
    hs.add_dev(wdev, segs, odev, 16)
    probes = [(4, np.zeros((0, 4), np.uint32)), (3, items[3][:1]), (5, np.concatenate([items[5][::3], items[5][:5] ^ 1]))]
    fwords, fsegs, fn = svdag_py.make_segments(probes)
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    ref_dev = svdag_py.DeviceBuffer(4 * fn)
    rdev = svdag_py.DeviceBuffer(4 * fn)
    hs.find_dev(fdev, fsegs, ref_dev, 0)
    hs.find_dev(fdev, fsegs, rdev, variant)
    ref = ref_dev.to_array(np.uint32, fn)
    assert np.array_equal(rdev.to_array(np.uint32, fn), ref)
    assert (ref[:1 + len(items[5][::3])] != svdag_py.NOT_FOUND).all()
    assert (ref[1 + len(items[5][::3]):] == svdag_py.NOT_FOUND).all()
    hs.close()
