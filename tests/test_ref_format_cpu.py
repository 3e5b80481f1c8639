"""The oracle's restatement of the reference's version-3 file format
(oracle/ref_format.py), on the CPU.

No file the reference wrote is available (its scene caches are not in the
repository) and its serialisers need CUDA to build, so these checks pin the
restatement by round trips and by the invariants of the reference's table
layout rather than by reference-written bytes:
  - a DAG written and read back has the oracle DAG's canonical hash, at
    material bits 0, 4 and 8;
  - the fully-filled chain is the full leaf of each material and, above it,
    the full node of the level below, as MyGPUHashDAG::allocate stores it;
  - every stored item sits in the chain of its bucket (xorHash32xN % buckets)
    with its acceleration byte, every overflow slab's allocator bit is set and
    no other bit is, and the per-table counts agree with the slabs;
  - an Engine file with voxel textures parses to its end.
"""
import struct

import numpy as np
import pytest

import oracle_py
import ref_format
import workloads


def _oracle(levels, M, edits=True):
    tp = workloads.terrain_params(levels)
    o = oracle_py.OracleDag(levels, M)
    mat = min(tp["material"], (1 << M) - 1) if M else 0
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], mat)
    if edits:
        side = float(1 << levels)
        o.edit(0, (3 % (1 << M)) if M else 0, (side * 0.3, side * 0.55, side * 0.4), side * 0.12)
        o.edit(1, 0, (side * 0.6, side * 0.45, side * 0.5), side * 0.1)
        if M:
            o.edit(4, 5 % (1 << M), (side * 0.5, side * 0.45, side * 0.5), side * 0.2)
    return o


@pytest.mark.parametrize("M", [0, 4, 8])
def test_write_then_read_gives_the_same_dag(M):
    levels = 8
    o = _oracle(levels, M)
    h, counts, dump = o.canonical(dump=True)
    data = ref_format.write_dag(levels, M, dump)
    lv, m2, root, recs, used, ff = ref_format.read_dag(data)
    assert (lv, m2, used) == (levels, M, len(data))
    h2, counts2 = oracle_py.canonical_from_records(recs)
    assert h2 == h and counts2 == counts


def _tables(data):
    """(levels, M, ff, tables) with each table's (s, buckets, heads, supers)."""
    r = ref_format._In(data)
    r.get("i")
    n, _ = r.get("Qi")
    ff = r.words(n).copy()
    levels, mats = r.get("II")
    r.get("I")
    nt, _ = r.get("Qi")
    out = []
    for _i in range(nt):
        s, sw, nb = r.get("III")
        nh, _ = r.get("Ii")
        heads = r.words(nh)
        r.get("II")
        nsup, _ = r.get("Qi")
        supers = []
        for _k in range(nsup):
            k, _ = r.get("Qi")
            supers.append(r.words(k))
        alloc_cap = r.get("Q")
        units = r.get("I")
        n_items, n_slabs = r.get("II")
        out.append(dict(s=s, sw=sw, nb=nb, heads=heads, supers=supers, cap=alloc_cap, units=units, items=n_items,
                        slabs=n_slabs))
    assert r.p == len(data)
    return levels, {1: 0, 16: 4, 256: 8}[mats], ff, out


@pytest.mark.parametrize("M", [0, 4])
def test_fully_filled_chain_is_stored_as_allocate_stores_it(M):
    levels = 7
    o = _oracle(levels, M, edits=False)
    _, _, dump = o.canonical(dump=True)
    data = ref_format.write_dag(levels, M, dump)
    _, _, ff, tables = _tables(data)
    leaf_level = levels - 2

    def item(leaf, h):
        s, ptr = ref_format.decode_handle(M, leaf, int(h))
        t = tables[s - 2]
        slot, head, slab = ptr & 31, (ptr >> 5) & 1, ptr >> 6
        if head:
            sl = t["heads"][slab * t["sw"]:(slab + 1) * t["sw"]]
        else:
            base = ref_format.BITMAP_WORDS + (slab & 16383) * t["sw"]
            sl = t["supers"][slab >> 14][base:base + t["sw"]]
        return [int(x) for x in sl[10 + slot * s:10 + (slot + 1) * s]]

    for m in range(1 << M):
        full = oracle_py.leaf_build(M, [1] * 64, [m] * 64)
        assert item(True, ff[m * levels + leaf_level]) == [int(x) for x in full]
        assert all(oracle_py.leaf_material(M, full, v) == m for v in range(64))
        for L in range(leaf_level - 1, -1, -1):
            w = item(False, ff[m * levels + L])
            assert len(w) == 9 and w[1:] == [int(ff[m * levels + L + 1])] * 8
            assert oracle_py.header_decode(M, w[0]) == (0xFF, 0, 1, m)  # mask, payload, full, material


def test_items_sit_in_their_buckets_and_allocator_bits_match():
    levels, M = 8, 4
    o = _oracle(levels, M)
    _, _, dump = o.canonical(dump=True)
    data = ref_format.write_dag(levels, M, dump)
    _, _, _, tables = _tables(data)
    for t in tables:
        s, sw, nb = t["s"], t["sw"], t["nb"]
        assert sw == 10 + 32 * s and len(t["heads"]) == nb * sw
        units = t["units"]
        seen_units = set()
        stored = 0
        for b in range(nb):
            sl = t["heads"][b * sw:(b + 1) * sw]
            while True:
                tags = np.frombuffer(sl[:8].tobytes(), np.uint8)
                for i in np.nonzero(tags)[0]:
                    w = sl[10 + i * s:10 + (i + 1) * s]
                    assert oracle_py.xor_hash(w) % nb == b
                    assert oracle_py.accel_hash(w) == tags[i]
                    stored += 1
                nxt = int(sl[8])
                if nxt == 0xFFFFFFFF:
                    break
                assert nxt < units and nxt not in seen_units
                seen_units.add(nxt)
                sup = t["supers"][nxt >> 14]
                assert (int(sup[((nxt >> 10) & 15) * 32 + (nxt & 1023) // 32]) >> (nxt & 31)) & 1
                base = ref_format.BITMAP_WORDS + ((nxt >> 10) & 15) * 1024 * sw + (nxt & 1023) * sw
                sl = sup[base:base + sw]
        assert seen_units == set(range(units))
        bits = sum(bin(int(x)).count("1") for sup in t["supers"] for x in sup[:ref_format.BITMAP_WORDS])
        assert bits == units
        assert stored == t["items"] and t["slabs"] == nb + units and t["cap"] == len(t["supers"])
        assert len(t["supers"]) == max(-(-(nb + 1) // 16384), -(-units // 16384))


def test_items_are_stored_once():
    # Find-before-add keeps every content once per table, the fully-filled
    # chain included, even where the terrain holds full subtrees.
    levels, M = 8, 4
    o = _oracle(levels, M)
    _, _, dump = o.canonical(dump=True)
    _, _, _, tables = _tables(ref_format.write_dag(levels, M, dump))
    for t in tables:
        s, sw = t["s"], t["sw"]
        seen = set()
        slabs = [t["heads"][b * sw:(b + 1) * sw] for b in range(t["nb"])]
        for u in range(t["units"]):
            base = ref_format.BITMAP_WORDS + ((u >> 10) & 15) * 1024 * sw + (u & 1023) * sw
            slabs.append(t["supers"][u >> 14][base:base + sw])
        for sl in slabs:
            tags = np.frombuffer(sl[:8].tobytes(), np.uint8)
            for i in np.nonzero(tags)[0]:
                key = tuple(int(x) for x in sl[10 + i * s:10 + (i + 1) * s])
                assert key not in seen
                seen.add(key)


def test_engine_file_with_textures_parses_to_its_end():
    levels, M = 7, 4
    o = _oracle(levels, M)
    h, _, dump = o.canonical(dump=True)
    img = (np.arange(4 * 3 * 2, dtype=np.uint8), 3, 2)
    mats = [("stone", [img, img, img], [0.5] * 9), ("grass", [img, img, img], [0.25] * 9)]
    bounds = (-1.0, -2.0, -3.0, 4.0, 5.0, 6.0)
    view = tuple(float(i) for i in range(12))
    data = ref_format.write_scene(levels, M, dump, bounds, view, mats)
    b2, recs, v2, names = ref_format.read_scene(data)
    assert b2 == bounds and v2 == view and names == ["stone", "grass"]
    assert oracle_py.canonical_from_records(recs)[0] == h
    # DAGInfo is 48 bytes of doubles ahead of the DAG's version word.
    assert struct.unpack_from("<i", data, 48)[0] == 3
