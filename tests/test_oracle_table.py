"""Pins the oracle's restatement of CompactAccelerationHashTable (host add/find)
against brute force and the published hash formulas."""
import numpy as np
import pytest

import oracle_py
import workloads


def py_xor_hash(item):
    # utils.h:386-403, evaluated with Python integers.
    seed = 0
    for i, w in enumerate(int(x) for x in item):
        seed ^= ((2717583853 * w + i * 1117415409) % (1 << 32)) % 4294967291
    return seed


def py_accel_hash(item):
    # compact_acceleration_hash_table.h:493-503.
    out = 0
    for i, w in enumerate(int(x) for x in item):
        out ^= ((32 - i) * w) % (1 << 32)
    out &= 0xFF
    return out or 1


def test_hash_formulas_against_python_integers():
    rng = np.random.default_rng(11)
    for s in range(1, 19):
        for _ in range(20):
            item = rng.integers(0, 1 << 32, s, dtype=np.uint64).astype(np.uint32)
            assert oracle_py.xor_hash(item) == py_xor_hash(item)
            assert oracle_py.accel_hash(item) == py_accel_hash(item)
    # Words at or above the prime wrap: 0xFFFFFFFF*X + 0 mod 2^32 exercised above;
    # a zero acceleration tag is remapped to 1.
    assert oracle_py.accel_hash(np.zeros(3, np.uint32)) == 1


@pytest.mark.parametrize("s,nb", [(2, 1), (2, 7), (5, 3), (9, 64), (18, 5)])
def test_table_matches_brute_force(s, nb):
    rng = np.random.default_rng(s * 100 + nb)
    n = 3000
    items = rng.integers(0, 50, size=(n, s)).astype(np.uint32)  # many duplicates
    t = oracle_py.OracleTable(s, nb)
    ptrs = t.add(items)
    # Every insert returns a pointer that decodes to the inserted words.
    assert np.array_equal(t.decode(ptrs), items)
    present = {tuple(x) for x in items.tolist()}
    probes = rng.integers(0, 60, size=(4000, s)).astype(np.uint32)
    found = t.find(probes)
    for p, f in zip(probes.tolist(), found.tolist()):
        assert (f != 0xFFFFFFFF) == (tuple(p) in present)
    hit = found != 0xFFFFFFFF
    assert np.array_equal(t.decode(found[hit]), probes[hit])
    # Threaded finds are the same as sequential ones.
    assert np.array_equal(t.find(probes, threads=4), found)


def test_chains_grow_past_one_slab():
    t = oracle_py.OracleTable(3, 1)
    items = np.arange(300 * 3, dtype=np.uint32).reshape(300, 3)
    t.add(items)
    assert t.overflow_slabs() == (300 - 32 + 31) // 32
    assert np.all(t.find(items) != 0xFFFFFFFF)


def test_workload_shape():
    items = workloads.slab_hash_items(1 << 14)
    assert sum(len(a) for a in items.values()) == 1 << 14
    for s, a in items.items():
        assert a.shape[1] == s
        assert np.all(workloads.POPC8[a[:, 0]] == s - 1)
    finds = workloads.slab_hash_finds(items, 1 << 14)
    assert abs(sum(len(a) for a in finds.values()) - (1 << 14)) <= 8
