"""ORACLE - test infrastructure only: tests/ use it as the checker of the
library's reference-format reader and writer; the product never imports it.

Restatement of the reference's version-3 on-disk format, written from its
serialisers:
  - BinaryWriter / BinaryReader (binary_writer.h:84-119, binary_reader.h:92-130):
    trivially copyable values raw (little endian), std::vector as u64 count
    plus elements, std::string as u64 length plus bytes;
  - StaticArray<T, Size> (array.h:30-82): Size count, int32 memory type, then
    the elements (raw, or each one's writeTo); DynamicArray adds its u64
    allocated count (array.h:311-320); StaticArray2D writes its buffer, then
    u32 width and height (array2d.h:113-123);
  - MyGPUHashDAG::writeTo (my_gpu_hash_dag.cu:157-165): int32 version 3,
    fullyFilledNodes (width = levels, height = materials, entry
    [material * levels + level]), u32 firstNodeIndex, hashTables
    (StaticArray<table>, one per item size 2 .. max(9, 2 + 2M));
  - GpuHashTableBase::writeTo (gpu_hash_table_base.cu:116-136): u32 item size,
    u32 slab size (10 + 32 s, compact_acceleration_hash_table.cu:5-13),
    u32 numBuckets, the head slabs (StaticArray<u32, u32>), the slab
    allocator, u32 numItems, u32 numSlabs;
  - DynamicSlabAllocator::writeTo (dynamic_slab_alloc.cu:172-179): u32 unit
    size, u32 super-block size (512 + 16 * 1024 * unit size,
    dynamic_slab_alloc.cu:33-37 and dynamic_slab_alloc.h:357-362), the super
    blocks (DynamicArray of StaticArray<u32, u64>, each 512 bitmap words then
    the units), u32 allocated units;
  - Engine::writeTo (engine.cpp:1551-1557): DAGInfo (6 doubles, dag_info.h),
    the DAG, CameraView (12 doubles, camera_view.h:12-13), VoxelTextures
    (voxel_textures.cpp:352-383, image.h:52-68).

Slabs, pointers and hashes are OracleTable's (oracle_table.cpp, the
restatement of CompactAccelerationHashTable).  An overflow slab's allocator
pointer is super block << 14 | block << 10 | unit (dynamic_slab_alloc.h:254-265,
357-370); units are handed out in order here, so overflow slab k is pointer k.
Handles (my_gpu_base_dag.h:209-238, 334-342): node = pointer | (s - 2) << 29;
leaf = pointer | (s - minLeaf) << (32 - b), b = bit_width(maxLeaf - minLeaf).
"""
import struct

import numpy as np

import oracle_py

VERSION = 3
MEM_GPU_MALLOC = 1
MEM_CPU = 3
UNITS_PER_SUPER = 16 * 1024
BITMAP_WORDS = 16 * 1024 // 32
LOAD_FACTOR = 96


def slab_words(s):
    return 10 + 32 * s


def super_words(s):
    return BITMAP_WORDS + UNITS_PER_SUPER * slab_words(s)


def min_leaf(M):
    return 3 if M else 2


def max_leaf(M):
    return 2 + (64 * M + 31) // 32


def num_tables(M):
    return max(9, max_leaf(M)) - 2 + 1


def leaf_bits(M):
    return (max_leaf(M) - min_leaf(M)).bit_length()


def node_handle(ptr, s):
    assert ptr < (1 << 29)
    return ptr | ((s - 2) << 29)


def leaf_handle(M, ptr, s):
    b = leaf_bits(M)
    if b == 0:
        return ptr
    assert ptr < (1 << (32 - b))
    return ptr | ((s - min_leaf(M)) << (32 - b))


def decode_handle(M, leaf, h):
    """(item size, table pointer) of a handle."""
    if not leaf:
        return 2 + (h >> 29), h & ((1 << 29) - 1)
    b = leaf_bits(M)
    if b == 0:
        return min_leaf(M), h
    return min_leaf(M) + (h >> (32 - b)), h & ((1 << (32 - b)) - 1)


def full_node_header(M, m):
    """encode_node_header(p, 0xFF, material) (my_gpu_base_dag.h:167-175, 343-350)."""
    return 0xFF | (1 << 8) | ((m << 9) if M else 0)


def parse_dump(dump, leaf_level):
    """Canonical dump -> ({level: [item words]}, root id)."""
    d = [int(x) for x in dump]
    p, out = 0, {}
    for L in range(leaf_level, -1, -1):
        assert d[p] == L
        count = d[p + 1]
        p += 2
        items = []
        for _ in range(count):
            s = d[p]
            items.append(d[p + 1:p + 1 + s])
            p += 1 + s
        out[L] = items
    assert p + 1 == len(d)
    return out, d[p]


class _Out:
    def __init__(self):
        self.parts = []

    def put(self, fmt, *v):
        self.parts.append(struct.pack("<" + fmt, *v))

    def words(self, a):
        self.parts.append(np.ascontiguousarray(a, np.uint32).tobytes())

    def bytes(self):
        return b"".join(self.parts)


def write_dag(levels, M, dump):
    """MyGPUHashDAG::writeTo bytes of the canonical dump's DAG."""
    leaf_level = levels - 2
    mats = 1 << M
    items, root_id = parse_dump(dump, leaf_level)
    nt = num_tables(M)
    per = [0] * nt
    for L in items:
        for w in items[L]:
            per[len(w) - 2] += 1
    per[max_leaf(M) - 2] += mats
    per[9 - 2] += mats * leaf_level
    buckets = [n // LOAD_FACTOR + 1 for n in per]
    tables = [oracle_py.OracleTable(2 + i, buckets[i]) for i in range(nt)]

    def store(leaf, w):
        s = len(w)
        t = tables[s - 2]
        a = np.array(w, np.uint32)
        ptr = int(t.find(a)[0])
        if ptr == 0xFFFFFFFF:
            ptr = int(t.add(a)[0])
        return leaf_handle(M, ptr, s) if leaf else node_handle(ptr, s)

    ff = [0xFFFFFFFF] * (levels * mats)
    for m in range(mats):
        full = oracle_py.leaf_build(M, np.ones(64, np.uint32), np.full(64, m, np.uint32))
        ff[m * levels + leaf_level] = store(True, [int(x) for x in full])
    for L in range(leaf_level - 1, -1, -1):
        for m in range(mats):
            ff[m * levels + L] = store(False, [full_node_header(M, m)] + [ff[m * levels + L + 1]] * 8)
    below = []
    for L in range(leaf_level, -1, -1):
        leaf = L == leaf_level
        below = [store(leaf, w if leaf else [w[0]] + [below[c] for c in w[1:]]) for w in items[L]]
    root = 0xFFFFFFFF if root_id == 0xFFFFFFFF else below[root_id]

    o = _Out()
    o.put("i", VERSION)
    o.put("Qi", len(ff), MEM_GPU_MALLOC)
    o.words(ff)
    o.put("II", levels, mats)
    o.put("I", root)
    o.put("Qi", nt, MEM_GPU_MALLOC)
    for i, t in enumerate(tables):
        s = 2 + i
        sw = slab_words(s)
        heads, pool = t.raw(buckets[i])
        units = len(pool) // sw
        o.put("III", s, sw, buckets[i])
        o.put("Ii", len(heads), MEM_GPU_MALLOC)
        o.words(heads)
        nsup = max(-(-(buckets[i] + 1) // UNITS_PER_SUPER), -(-units // UNITS_PER_SUPER))
        o.put("II", sw, super_words(s))
        o.put("Qi", nsup, MEM_CPU)
        for k in range(nsup):
            blk = np.zeros(super_words(s), np.uint32)
            used = max(0, min(UNITS_PER_SUPER, units - k * UNITS_PER_SUPER))
            for u in range(used):
                blk[u // 32] |= np.uint32(1 << (u % 32))
            blk[BITMAP_WORDS:BITMAP_WORDS + used * sw] = pool[k * UNITS_PER_SUPER * sw:(k * UNITS_PER_SUPER + used) * sw]
            o.put("Qi", len(blk), MEM_GPU_MALLOC)
            o.words(blk)
        o.put("Q", nsup)
        o.put("I", units)
        n_items = sum(int(np.count_nonzero(np.frombuffer(a[j:j + 8].tobytes(), np.uint8)))
                      for a in (heads, pool) for j in range(0, len(a), sw))
        o.put("II", n_items, buckets[i] + units)
    return o.bytes()


class _In:
    def __init__(self, data, off=0):
        self.d = memoryview(data)
        self.p = off

    def get(self, fmt):
        v = struct.unpack_from("<" + fmt, self.d, self.p)
        self.p += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def words(self, n):
        a = np.frombuffer(self.d, np.uint32, n, self.p)
        self.p += 4 * n
        return a


def read_dag(data, off=0):
    """Parses one MyGPUHashDAG stream; returns (levels, M, root handle,
    export records of every (level, handle) reachable from the root, bytes
    read).  The records' layout is the library's svdag_export_tables layout,
    so oracle_py.canonical_from_records gives the canonical hash."""
    r = _In(data, off)
    assert r.get("i") == VERSION
    nff, _ = r.get("Qi")
    ff = r.words(nff)
    levels, mats = r.get("II")
    M = {1: 0, 16: 4, 256: 8}[mats]
    root = r.get("I")
    nt, _ = r.get("Qi")
    assert nt == num_tables(M)
    tables = []
    for i in range(nt):
        s, sw, nb = r.get("III")
        assert s == 2 + i and sw == slab_words(s)
        nh, _ = r.get("Ii")
        heads = r.words(nh)
        unit, sup = r.get("II")
        assert unit == sw and sup == super_words(s)
        nsup, _ = r.get("Qi")
        supers = []
        for _k in range(nsup):
            n, _ = r.get("Qi")
            supers.append(r.words(n))
        r.get("Q")
        r.get("I")
        r.get("II")
        tables.append((s, sw, nb, heads, supers))

    def item(leaf, h):
        s, ptr = decode_handle(M, leaf, h)
        _s, sw, nb, heads, supers = tables[s - 2]
        slot, head, slab = ptr & 31, (ptr >> 5) & 1, ptr >> 6
        if head:
            assert slab < nb
            sl = heads[slab * sw:(slab + 1) * sw]
        else:
            sb, blk, u = slab >> 14, (slab >> 10) & 15, slab & 1023
            assert (int(supers[sb][blk * 32 + u // 32]) >> (u % 32)) & 1
            base = BITMAP_WORDS + (blk * 1024 + u) * sw
            sl = supers[sb][base:base + sw]
        assert (int(sl[slot // 4]) >> (8 * (slot % 4))) & 0xFF
        return [int(x) for x in sl[10 + slot * s:10 + (slot + 1) * s]]

    leaf_level = levels - 2
    recs = []
    cur = [root] if root != 0xFFFFFFFF else []
    for L in range(leaf_level + 1):
        nxt = []
        for h in sorted(set(cur)):
            w = item(L == leaf_level, h)
            recs += [L, h, len(w)] + w
            if L < leaf_level:
                nxt += w[1:]
        cur = nxt
    n = 0
    p = 0
    while p < len(recs):
        n += 1
        p += 3 + recs[p + 2]
    header = [levels, M, root, n]
    return levels, M, root, np.array(header + recs, np.uint32), r.p - off, ff


def write_scene(levels, M, dump, bounds=None, view=None, materials=()):
    """Engine::writeTo bytes; materials: [(name, [3 x (pixels u8 array of
    4 * w * h, w, h)], 9 floats of averages)] for the VoxelTextures part."""
    side = float(1 << levels)
    bounds = bounds or (0.0, 0.0, 0.0, side, side, side)
    view = view or (0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1)
    o = _Out()
    o.put("6d", *bounds)
    o.parts.append(write_dag(levels, M, dump))
    o.put("12d", *[float(x) for x in view])
    o.put("Q", len(materials))
    for name, _imgs, _avg in materials:
        b = name.encode()
        o.put("Q", len(b))
        o.parts.append(b)
    o.put("Q", len(materials))
    for _name, imgs, avg in materials:
        for px, w, h in imgs:
            px = np.ascontiguousarray(px, np.uint8)
            o.put("Q", len(px) // 4)
            o.parts.append(px.tobytes())
            o.put("iiff?", w, h, float(w), float(h), False)
        o.put("9f", *avg)
    return o.bytes()


def read_scene(data):
    """(bounds, records, view, material names) of an Engine::writeTo file."""
    bounds = struct.unpack_from("<6d", data, 0)
    levels, M, root, recs, used, _ff = read_dag(data, 48)
    r = _In(data, 48 + used)
    view = r.get("12d")
    nn = r.get("Q")
    names = []
    for _ in range(nn):
        ln = r.get("Q")
        names.append(bytes(r.d[r.p:r.p + ln]).decode())
        r.p += ln
    nm = r.get("Q")
    for _ in range(nm):
        for _img in range(3):
            npx = r.get("Q")
            r.p += 4 * npx
            r.get("iiff?")
        r.get("9f")
    assert r.p == len(data)
    return bounds, recs, view, names
