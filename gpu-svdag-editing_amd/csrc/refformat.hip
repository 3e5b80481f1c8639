// The reference's on-disk format, version 3 (host side): read a DAG that
// MyGPUHashDAG::writeTo / Engine::writeTo wrote, and write one that their
// readFrom can load.
//
// Byte stream (BinaryWriter, binary_writer.h:84-119: trivially copyable
// values raw in the host's little-endian order; StaticArray<T, Size> as
// [Size count][int32 memory type][elements], array.h:30-82; DynamicArray adds
// its u64 allocated count, array.h:311-320):
//   MyGPUHashDAG (my_gpu_hash_dag.cu:157-177)
//     int32 version = 3
//     fullyFilledNodes: StaticArray<u32, u64> [material * levels + level],
//                       u32 width = levels, u32 height = 2^M   (array2d.h:113-123)
//     u32 firstNodeIndex (the root handle)
//     hashTables: StaticArray<table, u64>, one per item size 2 .. max(9, 2 + 2M)
//   table (gpu_hash_table_base.cu:116-162)
//     u32 itemSizeInU32 s, u32 slabSizeInU32 (10 + 32 s), u32 numBuckets,
//     StaticArray<u32, u32> head slabs (numBuckets * slab size words),
//     slab allocator (dynamic_slab_alloc.cu:172-188): u32 unit size (= slab
//       size), u32 super-block size, DynamicArray of super blocks (each a
//       StaticArray<u32, u64> of 512 bitmap words then 16 x 1024 units),
//       u32 allocated units,
//     u32 numItems, u32 numSlabs
//   Engine (engine.cpp:1543-1557): DAGInfo (two Vector3 of doubles), the DAG,
//     CameraView (Vector3 position, Matrix3x3 rotation, doubles), and
//     VoxelTextures (vector<string> names, vector<VoxelTexturesCPU>).
//
// Slab (compact_acceleration_hash_table.h:519-541): words 0-7 hold 32 one-byte
// acceleration hashes (slot i in byte i % 4 of word i / 4, 0 = empty), word 8
// the next slab (0xFFFFFFFF ends the chain), word 9 the active mask, and
// item i at word 10 + i * s.  Table pointer (gpu_hash_table_base.h:301-310):
// slot in bits 0-4, bit 5 set for a head slab, bits 6-31 the bucket (head
// slab) or the allocator pointer (super block << 14 | block << 10 | unit,
// dynamic_slab_alloc.h:357-370).  Handles keep the item size in their top
// bits (my_gpu_base_dag.h:209-238, 334-342): 3 bits for nodes, bit_width(2M)
// for leaves (none at M = 0).
#include "dag_internal.h"
#include "host_threads.h"
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

namespace svdag {
namespace {

constexpr int32_t kRefVersion = 3;
constexpr uint32_t kRefSlots = 32;
constexpr uint32_t kRefNextWord = 8;
constexpr uint32_t kRefItemWord = 10;
constexpr uint32_t kRefNoSlab = 0xFFFFFFFFu;
constexpr uint32_t kRefUnitsPerBlock = 1024;
constexpr uint32_t kRefBlocksPerSuper = 16;
constexpr uint32_t kRefBitmapWords = kRefBlocksPerSuper * kRefUnitsPerBlock / 32; // 512
constexpr uint32_t kRefUnitsPerSuper = kRefBlocksPerSuper * kRefUnitsPerBlock;    // 16384
constexpr uint32_t kRefSlabFieldBits = 26;
constexpr uint32_t kRefLoadFactor = 96; // TARGET_LOAD_FACTOR (gpu_hash_dag_definitions.h:4)
// EMemoryType (memory.h:17-23): what a GPU_Malloc DAG's arrays carry, and the
// allocator's host-side list of super blocks.
constexpr int32_t kMemGpuMalloc = 1;
constexpr int32_t kMemCpu = 3;
constexpr int32_t kMemUnknown = 4;

inline uint32_t ref_slab_words(uint32_t s) { return kRefItemWord + kRefSlots * s; }
inline uint64_t ref_super_words(uint32_t s) { return kRefBitmapWords + (uint64_t)kRefUnitsPerSuper * ref_slab_words(s); }
inline uint32_t ref_num_tables(uint32_t M) { return std::max(9u, 2u + 2u * M) - 1u; }

// computeAccelerationHash (compact_acceleration_hash_table.h:491-501).
inline uint32_t ref_accel(const uint32_t* w, uint32_t s)
{
    uint32_t h = 0;
    for (uint32_t i = 0; i < s; ++i)
        h ^= (32u - i) * w[i];
    h &= 0xFFu;
    return h ? h : 1u;
}
// computeBucket with SlabHashXor (gpu_hash_table_base.h:268-284, utils.h:395-403).
inline uint32_t ref_bucket_hash(const uint32_t* w, uint32_t s)
{
    uint32_t h = 0;
    for (uint32_t j = 0; j < s; ++j)
        h ^= hash_part(w[j], j);
    return h;
}

inline uint32_t ld32(const uint8_t* base, uint64_t word)
{
    uint32_t v;
    std::memcpy(&v, base + word * 4, 4);
    return v;
}

struct ByteReader {
    const uint8_t* p;
    size_t n, off = 0;
    bool ok = true;
    template <typename T>
    T get()
    {
        T v {};
        if (!ok || n - off < sizeof(T)) {
            ok = false;
            return v;
        }
        std::memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
    const uint8_t* take(uint64_t bytes)
    {
        if (!ok || bytes > n - off) {
            ok = false;
            return nullptr;
        }
        const uint8_t* q = p + off;
        off += bytes;
        return q;
    }
    bool mem_type()
    {
        const int32_t t = get<int32_t>();
        ok = ok && t >= 0 && t < kMemUnknown;
        return ok;
    }
};

struct RefTableView {
    uint32_t s = 0, slabWords = 0, numBuckets = 0;
    const uint8_t* heads = nullptr;
    std::vector<const uint8_t*> supers;
};

struct RefDagView {
    uint32_t levels = 0, materials = 0, M = 0, root = kSentinel;
    std::vector<uint32_t> fullyFilled; // [material * levels + level]
    std::vector<RefTableView> tables;
};

int fail(const std::string& msg)
{
    set_error(msg);
    return SVDAG_E_INVALID;
}

// MyGPUHashDAG::readFrom's walk over the stream (my_gpu_hash_dag.cu:167-177),
// keeping views into the bytes instead of copies.
int parse_dag(ByteReader& r, RefDagView& v)
{
    const int32_t version = r.get<int32_t>();
    if (!r.ok || version != kRefVersion)
        return fail("reference DAG: not a version-3 MyGPUHashDAG stream (version " + std::to_string(version) + ")");
    const uint64_t ffCount = r.get<uint64_t>();
    if (!r.mem_type() || ffCount > (1u << 16))
        return fail("reference DAG: bad fullyFilledNodes array");
    const uint8_t* ff = r.take(ffCount * 4);
    v.levels = r.get<uint32_t>();
    v.materials = r.get<uint32_t>();
    if (!r.ok || (uint64_t)v.levels * v.materials != ffCount)
        return fail("reference DAG: fullyFilledNodes is not levels x materials");
    if (v.materials == 1)
        v.M = 0;
    else if (v.materials == 16)
        v.M = 4;
    else if (v.materials == 256)
        v.M = 8;
    else
        return fail("reference DAG: " + std::to_string(v.materials) + " materials (expected 1, 16 or 256)");
    v.fullyFilled.resize(ffCount);
    for (uint64_t i = 0; i < ffCount; ++i)
        v.fullyFilled[i] = ld32(ff, i);
    v.root = r.get<uint32_t>();
    const uint64_t nTables = r.get<uint64_t>();
    if (!r.mem_type() || nTables != ref_num_tables(v.M))
        return fail("reference DAG: " + std::to_string(nTables) + " hash tables, expected "
                    + std::to_string(ref_num_tables(v.M)) + " for " + std::to_string(v.M) + " material bits");
    v.tables.resize(nTables);
    for (uint32_t i = 0; i < nTables; ++i) {
        RefTableView& t = v.tables[i];
        t.s = r.get<uint32_t>();
        t.slabWords = r.get<uint32_t>();
        t.numBuckets = r.get<uint32_t>();
        const uint32_t headWords = r.get<uint32_t>();
        if (!r.mem_type() || t.s != 2 + i || t.slabWords != ref_slab_words(t.s) || t.numBuckets == 0
            || (uint64_t)t.numBuckets * t.slabWords != headWords || t.numBuckets >= (1u << kRefSlabFieldBits))
            return fail("reference DAG: table " + std::to_string(i) + " header does not describe a 32-slot slab table of "
                        + std::to_string(2 + i) + "-word items");
        t.heads = r.take((uint64_t)headWords * 4);
        const uint32_t unitWords = r.get<uint32_t>();
        const uint32_t superWords = r.get<uint32_t>();
        const uint64_t nSupers = r.get<uint64_t>();
        if (!r.mem_type() || unitWords != t.slabWords || superWords != ref_super_words(t.s)
            || nSupers > (1u << (kRefSlabFieldBits - 14)))
            return fail("reference DAG: table " + std::to_string(i) + " slab allocator header is inconsistent");
        for (uint64_t k = 0; k < nSupers; ++k) {
            const uint64_t words = r.get<uint64_t>();
            if (!r.mem_type() || words != superWords)
                return fail("reference DAG: table " + std::to_string(i) + " super block " + std::to_string(k) + " has the wrong size");
            t.supers.push_back(r.take(words * 4));
        }
        r.get<uint64_t>(); // allocated capacity of the super-block list
        r.get<uint32_t>(); // allocated units
        r.get<uint32_t>(); // numItems (statistics)
        r.get<uint32_t>(); // numSlabs (statistics)
        if (!r.ok)
            return fail("reference DAG: truncated in table " + std::to_string(i));
    }
    return SVDAG_OK;
}

// The words of the item a table pointer addresses (decodePointerImpl,
// gpu_hash_table_base.h:286-300, and DynamicSlabAllocator::decodePointer,
// dynamic_slab_alloc.h:231-238), or nullptr if the pointer leaves the table,
// names a unit the allocator's bitmap does not hold, or an empty slot.
const uint8_t* ref_item(const RefTableView& t, uint32_t ptr)
{
    const uint32_t slot = ptr & 31u, head = (ptr >> 5) & 1u, slab = ptr >> 6;
    const uint8_t* sp;
    if (head) {
        if (slab >= t.numBuckets)
            return nullptr;
        sp = t.heads + (uint64_t)slab * t.slabWords * 4;
    } else {
        const uint32_t sb = slab >> 14, blk = (slab >> 10) & 15u, unit = slab & 1023u;
        if (sb >= t.supers.size())
            return nullptr;
        const uint8_t* base = t.supers[sb];
        if (!((ld32(base, blk * 32u + unit / 32u) >> (unit % 32u)) & 1u))
            return nullptr;
        sp = base + (kRefBitmapWords + (uint64_t)(blk * kRefUnitsPerBlock + unit) * t.slabWords) * 4;
    }
    if (!((ld32(sp, slot / 4u) >> (8u * (slot % 4u))) & 0xFFu))
        return nullptr;
    return sp + ((uint64_t)kRefItemWord + (uint64_t)slot * t.s) * 4;
}

struct RefCodec {
    uint32_t minLeaf, leafSizeBits;
    explicit RefCodec(uint32_t M)
    {
        minLeaf = M ? 3u : 2u;
        const uint32_t span = 2u * M + 2u - minLeaf; // maxLeaf - minLeaf
        leafSizeBits = span ? 32u - (uint32_t)__builtin_clz(span) : 0u;
    }
    void decode(bool leaf, uint32_t h, uint32_t& s, uint32_t& ptr) const
    {
        if (leaf) {
            s = minLeaf + (leafSizeBits ? h >> (32u - leafSizeBits) : 0u);
            ptr = leafSizeBits ? h & ((1u << (32u - leafSizeBits)) - 1u) : h;
        } else {
            s = 2u + (h >> 29);
            ptr = h & ((1u << 29) - 1u);
        }
    }
    // kSentinel if ptr does not fit the handle's index field.
    uint32_t encode(bool leaf, uint32_t s, uint32_t ptr) const
    {
        if (leaf) {
            if (!leafSizeBits)
                return ptr == kSentinel ? kSentinel : ptr;
            if (ptr >> (32u - leafSizeBits))
                return kSentinel;
            return ptr | ((s - minLeaf) << (32u - leafSizeBits));
        }
        if (ptr >> 29)
            return kSentinel;
        return ptr | ((s - 2u) << 29);
    }
};

// The full leaf of one material: every voxel set, every material field m.
std::vector<uint32_t> full_leaf(uint32_t M, uint32_t m)
{
    std::vector<uint32_t> w(2 + 2 * M, 0u);
    w[0] = w[1] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < 64 && M; ++i)
        w[2 + (i * M) / 32] |= m << ((i * M) % 32);
    return w;
}

// Reads one DAG stream into ctx: handles reachable from the root are
// gathered level by level, their contents re-expressed with children as
// indices into the level below, duplicates (equal contents reached through
// different handles) merged, and the result imported as a dump.
int read_dag(svdag_ctx* ctx, ByteReader& r)
{
    RefDagView v;
    SVDAG_TRY(parse_dag(r, v));
    const DagDev& dag = ctx->dag;
    if (v.levels != dag.levels || v.M != dag.M)
        return fail("reference DAG: depth " + std::to_string(v.levels) + " with " + std::to_string(v.M)
                    + " material bits, but the context has depth " + std::to_string(dag.levels) + " with "
                    + std::to_string(dag.M));
    const RefCodec codec(v.M);
    const uint32_t leafLevel = dag.leafLevel;
    auto item = [&](bool leaf, uint32_t h, uint32_t& s) -> const uint8_t* {
        uint32_t ptr;
        codec.decode(leaf, h, s, ptr);
        if (s < 2 || s - 2 >= v.tables.size())
            return nullptr;
        return ref_item(v.tables[s - 2], ptr);
    };
    // The fully-filled chain must be what MyGPUHashDAG::allocate stores
    // (my_gpu_hash_dag.cu:38-63): a check that the stream fits this codec.
    for (uint32_t m = 0; m < v.materials; ++m) {
        uint32_t s;
        const uint8_t* p = item(true, v.fullyFilled[m * v.levels + leafLevel], s);
        const std::vector<uint32_t> want = full_leaf(v.M, m);
        bool same = p && s == want.size();
        for (uint32_t k = 0; same && k < s; ++k)
            same = ld32(p, k) == want[k];
        if (!same)
            return fail("reference DAG: fully-filled leaf of material " + std::to_string(m) + " is not the full leaf");
    }

    const uint32_t T = export_threads();
    std::vector<std::vector<uint32_t>> H(leafLevel + 1);
    if (v.root != kSentinel)
        H[0].push_back(v.root);
    std::string err;
    std::mutex errMu;
    auto note = [&](const std::string& e) {
        std::lock_guard<std::mutex> g(errMu);
        if (err.empty())
            err = e;
    };
    // Top down: every level's distinct handles.
    for (uint32_t L = 0; L < leafLevel; ++L) {
        const std::vector<uint32_t>& cur = H[L];
        std::vector<std::vector<uint32_t>> part(T);
        parallel_chunks(cur.size(), T, [&](size_t b, size_t e, uint32_t t) {
            for (size_t i = b; i < e; ++i) {
                uint32_t s;
                const uint8_t* p = item(false, cur[i], s);
                if (!p) {
                    note("reference DAG: a level-" + std::to_string(L) + " handle does not address a stored node");
                    return;
                }
                const uint32_t hdr = ld32(p, 0), mask = hdr & 0xFFu;
                if (!mask || s != 1u + (uint32_t)__builtin_popcount(mask)) {
                    note("reference DAG: a level-" + std::to_string(L) + " node's size does not match its child mask");
                    return;
                }
                for (uint32_t j = 1; j < s; ++j)
                    part[t].push_back(ld32(p, j));
            }
        });
        if (!err.empty())
            return fail(err);
        std::vector<uint32_t>& next = H[L + 1];
        size_t total = 0;
        for (auto& q : part)
            total += q.size();
        next.reserve(total);
        for (auto& q : part) {
            next.insert(next.end(), q.begin(), q.end());
            std::vector<uint32_t>().swap(q);
        }
        parallel_sort(next, T, std::less<uint32_t>());
        next.erase(std::unique(next.begin(), next.end()), next.end());
    }

    // Bottom up: contents with children as ids of the level below, merged.
    std::vector<uint32_t> dump, idBelow;
    for (int L = (int)leafLevel; L >= 0; --L) {
        const std::vector<uint32_t>& hs = H[L];
        const bool leaf = (uint32_t)L == leafLevel;
        const size_t n = hs.size();
        std::vector<uint32_t> size(n);
        std::vector<uint64_t> off(n + 1, 0);
        parallel_chunks(n, T, [&](size_t b, size_t e, uint32_t) {
            for (size_t i = b; i < e; ++i) {
                uint32_t s;
                if (!item(leaf, hs[i], s)) {
                    note("reference DAG: a level-" + std::to_string(L) + " handle does not address a stored item");
                    return;
                }
                size[i] = s;
            }
        });
        if (!err.empty())
            return fail(err);
        for (size_t i = 0; i < n; ++i)
            off[i + 1] = off[i] + size[i];
        std::vector<uint32_t> words(off[n]);
        std::vector<uint64_t> key(n);
        const std::vector<uint32_t>& below = L + 1 <= (int)leafLevel ? H[L + 1] : H[L];
        parallel_chunks(n, T, [&](size_t b, size_t e, uint32_t) {
            for (size_t i = b; i < e; ++i) {
                uint32_t s;
                const uint8_t* p = item(leaf, hs[i], s);
                uint32_t* w = words.data() + off[i];
                for (uint32_t k = 0; k < s; ++k)
                    w[k] = ld32(p, k);
                if (leaf) {
                    if (s != leaf_size_from_words(w[0], w[1], v.M) || (w[0] == 0 && w[1] == 0)) {
                        note("reference DAG: a leaf's size does not match its voxel mask");
                        return;
                    }
                } else {
                    for (uint32_t j = 1; j < s; ++j) {
                        const auto it = std::lower_bound(below.begin(), below.end(), w[j]);
                        w[j] = idBelow[(size_t)(it - below.begin())];
                    }
                }
                uint64_t x = 1469598103934665603ull ^ s;
                for (uint32_t k = 0; k < s; ++k)
                    x = (x ^ w[k]) * 1099511628211ull;
                key[i] = x;
            }
        });
        if (!err.empty())
            return fail(err);
        std::vector<uint32_t> order(n);
        for (size_t i = 0; i < n; ++i)
            order[i] = (uint32_t)i;
        auto less = [&](uint32_t a, uint32_t b) {
            if (key[a] != key[b])
                return key[a] < key[b];
            if (size[a] != size[b])
                return size[a] < size[b];
            return std::lexicographical_compare(words.begin() + (ptrdiff_t)off[a], words.begin() + (ptrdiff_t)off[a + 1],
                                                words.begin() + (ptrdiff_t)off[b], words.begin() + (ptrdiff_t)off[b + 1]);
        };
        parallel_sort(order, T, less);
        std::vector<uint32_t> id(n);
        const size_t head = dump.size();
        dump.push_back((uint32_t)L);
        dump.push_back(0u);
        uint32_t count = 0;
        for (size_t k = 0; k < n; ++k) {
            const uint32_t i = order[k];
            if (k == 0 || less(order[k - 1], i)) {
                dump.push_back(size[i]);
                dump.insert(dump.end(), words.begin() + (ptrdiff_t)off[i], words.begin() + (ptrdiff_t)off[i + 1]);
                ++count;
            }
            id[i] = count - 1;
        }
        dump[head + 1] = count;
        idBelow.swap(id);
    }
    dump.push_back(H[0].empty() ? kSentinel : idBelow[0]);
    return import_dump(ctx, dump.data(), dump.size());
}

// One table being built the way CompactAccelerationHashTable::add fills it
// (compact_acceleration_hash_table.h:27-60): the first empty slot along the
// bucket's chain; a full chain gets a new slab linked right after the head
// slab (allocateNewSlabThread, gpu_hash_table_base.h:120-160).  Allocator
// units are handed out in order, so unit n's allocator pointer is n.
struct RefTableBuild {
    uint32_t s = 0, slabWords = 0, numBuckets = 0, numItems = 0;
    std::vector<uint32_t> heads, units;
    uint32_t numUnits = 0;

    void init(uint32_t size, uint32_t buckets)
    {
        s = size;
        slabWords = ref_slab_words(size);
        numBuckets = buckets;
        heads.assign((size_t)buckets * slabWords, 0u);
        for (uint32_t b = 0; b < buckets; ++b)
            heads[(size_t)b * slabWords + kRefNextWord] = kRefNoSlab;
    }
    uint32_t* slab(uint32_t link) { return units.data() + (size_t)link * slabWords; }
    // Table pointer of the item (found or added); kSentinel if the pointer
    // would not fit its 26-bit slab field.
    uint32_t find_or_add(const uint32_t* w)
    {
        const uint32_t b = ref_bucket_hash(w, s) % numBuckets, a = ref_accel(w, s);
        uint32_t* head = heads.data() + (size_t)b * slabWords;
        auto ptr_of = [&](uint32_t link, uint32_t slot) {
            return link == kRefNoSlab ? (b << 6) | 32u | slot : (link << 6) | slot;
        };
        uint32_t link = kRefNoSlab;
        for (uint32_t* sp = head;;) {
            for (uint32_t i = 0; i < kRefSlots; ++i)
                if (((sp[i / 4] >> (8 * (i % 4))) & 0xFFu) == a
                    && std::equal(w, w + s, sp + kRefItemWord + (size_t)i * s))
                    return ptr_of(link, i);
            link = sp[kRefNextWord];
            if (link == kRefNoSlab)
                break;
            sp = slab(link);
        }
        link = kRefNoSlab;
        for (uint32_t* sp = head;;) {
            for (uint32_t i = 0; i < kRefSlots; ++i)
                if (!((sp[i / 4] >> (8 * (i % 4))) & 0xFFu)) {
                    sp[i / 4] |= a << (8 * (i % 4));
                    std::copy(w, w + s, sp + kRefItemWord + (size_t)i * s);
                    ++numItems;
                    return ptr_of(link, i);
                }
            link = sp[kRefNextWord];
            if (link == kRefNoSlab) {
                if (numUnits >= (1u << kRefSlabFieldBits))
                    return kSentinel;
                link = numUnits++;
                units.resize((size_t)numUnits * slabWords, 0u);
                head = heads.data() + (size_t)b * slabWords;
                uint32_t* fresh = slab(link);
                fresh[kRefNextWord] = head[kRefNextWord];
                head[kRefNextWord] = link;
                sp = fresh;
                continue;
            }
            sp = slab(link);
        }
    }
};

struct ByteWriter {
    std::vector<uint8_t> b;
    template <typename T>
    void put(T v)
    {
        const size_t o = b.size();
        b.resize(o + sizeof(T));
        std::memcpy(b.data() + o, &v, sizeof(T));
    }
    void words(const uint32_t* w, size_t n)
    {
        const size_t o = b.size();
        b.resize(o + n * 4);
        if (n)
            std::memcpy(b.data() + o, w, n * 4);
    }
    void zeros(size_t nWords) { b.resize(b.size() + nWords * 4, 0u); }
};

// Writes ctx's reachable DAG as MyGPUHashDAG::writeTo would for a GPU_Malloc
// DAG: MyGPUHashDAG::allocate's fully-filled chain first (find-or-add, as
// there), then the canonical dump's items from the leaves up, each found or
// added, with ~96 items per bucket (setLoadFactor(TARGET_LOAD_FACTOR),
// my_gpu_hash_dag.cu:139-155).  The caller holds ctx->mu.
int write_dag(svdag_ctx* ctx, ByteWriter& out)
{
    const DagDev& dag = ctx->dag;
    const uint32_t M = dag.M, levels = dag.levels, leafLevel = dag.leafLevel, materials = 1u << M;
    const RefCodec codec(M);
    std::vector<uint32_t> dump, counts;
    SVDAG_TRY(canonical_dump(ctx, dump, counts));
    // Item offsets per level (dump order: leaf level first).
    std::vector<std::vector<size_t>> at(leafLevel + 1);
    size_t p = 0;
    const uint32_t nTables = ref_num_tables(M);
    std::vector<uint64_t> perTable(nTables, 0);
    for (int L = (int)leafLevel; L >= 0; --L) {
        if (p + 2 > dump.size() || dump[p] != (uint32_t)L)
            return fail("reference write: malformed canonical dump");
        const uint32_t count = dump[p + 1];
        p += 2;
        at[L].resize(count);
        for (uint32_t i = 0; i < count; ++i) {
            at[L][i] = p;
            const uint32_t s = dump[p];
            if (s < 2 || s - 2 >= nTables)
                return fail("reference write: item size out of range");
            ++perTable[s - 2];
            p += 1 + s;
        }
    }
    const uint32_t rootId = dump[p];
    perTable[2u + 2u * M - 2u] += materials;        // full leaves
    perTable[9u - 2u] += (uint64_t)materials * leafLevel; // full nodes
    std::vector<RefTableBuild> tb(nTables);
    for (uint32_t i = 0; i < nTables; ++i) {
        const uint64_t buckets = perTable[i] / kRefLoadFactor + 1;
        if (buckets * ref_slab_words(2 + i) > 0xFFFFFFFFull)
            return fail("reference write: table of " + std::to_string(2 + i) + "-word items needs more head slabs than a u32 count holds");
        tb[i].init(2 + i, (uint32_t)buckets);
    }
    auto store = [&](bool leaf, const uint32_t* w, uint32_t s, uint32_t& handle) {
        const uint32_t ptr = tb[s - 2].find_or_add(w);
        handle = ptr == kSentinel ? kSentinel : codec.encode(leaf, s, ptr);
        if (handle == kSentinel) {
            set_error("reference write: the table of " + std::to_string(s)
                      + "-word items outgrows the reference's handle range");
            return SVDAG_E_CAPACITY;
        }
        return SVDAG_OK;
    };
    // fullyFilledNodes[material * levels + level]; levels below the leaf
    // level are never written by the reference (left as the sentinel here).
    std::vector<uint32_t> ff((size_t)materials * levels, kSentinel);
    for (uint32_t m = 0; m < materials; ++m) {
        const std::vector<uint32_t> w = full_leaf(M, m);
        SVDAG_TRY(store(true, w.data(), (uint32_t)w.size(), ff[(size_t)m * levels + leafLevel]));
    }
    for (int L = (int)leafLevel - 1; L >= 0; --L)
        for (uint32_t m = 0; m < materials; ++m) {
            uint32_t w[9];
            w[0] = header_make_full(0xFFu, m, M);
            for (uint32_t j = 1; j < 9; ++j)
                w[j] = ff[(size_t)m * levels + L + 1];
            SVDAG_TRY(store(false, w, 9, ff[(size_t)m * levels + L]));
        }
    // The reachable DAG, leaves first.  Tables are independent, so each
    // level's items go in on one thread per table, in level order within a
    // table: the same layout as one thread inserting them in order.
    std::vector<uint32_t> below, handles;
    std::vector<std::vector<uint32_t>> byTable(nTables);
    std::vector<uint32_t> failed(nTables);
    for (int L = (int)leafLevel; L >= 0; --L) {
        const bool leaf = (uint32_t)L == leafLevel;
        handles.assign(at[L].size(), kSentinel);
        for (auto& q : byTable)
            q.clear();
        for (size_t i = 0; i < at[L].size(); ++i)
            byTable[dump[at[L][i]] - 2].push_back((uint32_t)i);
        std::fill(failed.begin(), failed.end(), 0u);
        auto run = [&](uint32_t ti) {
            uint32_t w[kMaxItemWords];
            for (uint32_t i : byTable[ti]) {
                const uint32_t s = dump[at[L][i]];
                const uint32_t* src = dump.data() + at[L][i] + 1;
                w[0] = src[0];
                for (uint32_t j = 1; j < s; ++j)
                    w[j] = leaf ? src[j] : below[src[j]];
                const uint32_t ptr = tb[ti].find_or_add(w);
                handles[i] = ptr == kSentinel ? kSentinel : codec.encode(leaf, s, ptr);
                if (handles[i] == kSentinel) {
                    failed[ti] = 1;
                    return;
                }
            }
        };
        std::vector<std::thread> pool;
        for (uint32_t ti = 0; ti < nTables; ++ti)
            if (byTable[ti].size() >= 4096)
                pool.emplace_back(run, ti);
            else
                run(ti);
        for (auto& th : pool)
            th.join();
        for (uint32_t ti = 0; ti < nTables; ++ti)
            if (failed[ti]) {
                set_error("reference write: the table of " + std::to_string(ti + 2)
                          + "-word items outgrows the reference's handle range");
                return SVDAG_E_CAPACITY;
            }
        below.swap(handles);
    }
    const uint32_t root = rootId == kSentinel ? kSentinel : below[rootId];

    out.put<int32_t>(kRefVersion);
    out.put<uint64_t>(ff.size());
    out.put<int32_t>(kMemGpuMalloc);
    out.words(ff.data(), ff.size());
    out.put<uint32_t>(levels);
    out.put<uint32_t>(materials);
    out.put<uint32_t>(root);
    out.put<uint64_t>(nTables);
    out.put<int32_t>(kMemGpuMalloc);
    for (const RefTableBuild& t : tb) {
        out.put<uint32_t>(t.s);
        out.put<uint32_t>(t.slabWords);
        out.put<uint32_t>(t.numBuckets);
        out.put<uint32_t>((uint32_t)t.heads.size());
        out.put<int32_t>(kMemGpuMalloc);
        out.words(t.heads.data(), t.heads.size());
        // DynamicSlabAllocator::create reserves numBuckets + 1 units
        // (CompactAccelerationHashTable::allocate, compact_acceleration_hash_table.cu:5-12).
        const uint64_t superWords = ref_super_words(t.s);
        const uint64_t nSupers = std::max<uint64_t>((t.numBuckets + 1ull + kRefUnitsPerSuper - 1) / kRefUnitsPerSuper,
                                                    ((uint64_t)t.numUnits + kRefUnitsPerSuper - 1) / kRefUnitsPerSuper);
        out.put<uint32_t>(t.slabWords);
        out.put<uint32_t>((uint32_t)superWords);
        out.put<uint64_t>(nSupers);
        out.put<int32_t>(kMemCpu);
        for (uint64_t k = 0; k < nSupers; ++k) {
            out.put<uint64_t>(superWords);
            out.put<int32_t>(kMemGpuMalloc);
            const uint64_t first = k * kRefUnitsPerSuper;
            const uint64_t used = t.numUnits > first ? std::min<uint64_t>(kRefUnitsPerSuper, t.numUnits - first) : 0;
            uint32_t bitmap[kRefBitmapWords] = {};
            for (uint64_t u = 0; u < used; ++u)
                bitmap[u / 32] |= 1u << (u % 32);
            out.words(bitmap, kRefBitmapWords);
            out.words(t.units.data() + first * t.slabWords, used * t.slabWords);
            out.zeros((kRefUnitsPerSuper - used) * t.slabWords);
        }
        out.put<uint64_t>(nSupers);
        out.put<uint32_t>(t.numUnits);
        out.put<uint32_t>(t.numItems);
        out.put<uint32_t>(t.numBuckets + t.numUnits);
    }
    return SVDAG_OK;
}

// Engine::readFrom's trailing members (engine.cpp:1543-1549): CameraView,
// then VoxelTextures (voxel_textures.cpp:372-383): vector<string> names and
// vector<VoxelTexturesCPU>, each three Images (image.h:52-68: vector<Pixel>,
// int width, int height, float fWidth, float fHeight, bool alpha) and three
// float3 averages.  Parsed for their extent only.
bool skip_textures(ByteReader& r, uint64_t& numMaterials)
{
    const uint64_t names = r.get<uint64_t>();
    for (uint64_t i = 0; r.ok && i < names; ++i)
        r.take(r.get<uint64_t>());
    numMaterials = r.get<uint64_t>();
    for (uint64_t i = 0; r.ok && i < numMaterials; ++i) {
        for (int img = 0; r.ok && img < 3; ++img) {
            const uint64_t pixels = r.get<uint64_t>();
            if (pixels > (r.n - r.off) / 4)
                r.ok = false;
            r.take(pixels * 4);
            r.take(4 + 4 + 4 + 4 + 1);
        }
        r.take(3 * 12);
    }
    return r.ok;
}

} // namespace
} // namespace svdag

using namespace svdag;

extern "C" {

int svdag_read_reference_dag(svdag_ctx* ctx, const void* buf, size_t len, size_t* consumed)
{
    if (!ctx || (!buf && len))
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ByteReader r { (const uint8_t*)buf, len };
    SVDAG_TRY(read_dag(ctx, r));
    if (consumed)
        *consumed = r.off;
    return SVDAG_OK;
}

int svdag_write_reference_dag(svdag_ctx* ctx, void* buf, size_t* len)
{
    if (!ctx || !len)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    // A size query keeps its stream for the call that fetches it: the DAG
    // is the same while the root and the collection count are.
    const bool kept = !ctx->refStream.empty() && ctx->refStreamRoot == ctx->root && ctx->refStreamGc == ctx->gcSerial;
    if (!kept) {
        ByteWriter w;
        ctx->refStream.clear();
        SVDAG_TRY(write_dag(ctx, w));
        ctx->refStream.swap(w.b);
        ctx->refStreamRoot = ctx->root;
        ctx->refStreamGc = ctx->gcSerial;
    }
    const size_t n = ctx->refStream.size();
    if (buf && *len >= n) {
        std::memcpy(buf, ctx->refStream.data(), n);
        std::vector<uint8_t>().swap(ctx->refStream);
    }
    *len = n;
    return SVDAG_OK;
}

int svdag_read_reference_scene(svdag_ctx* ctx, const char* path, double* bounds, double* view)
{
    if (!ctx || !path)
        return SVDAG_E_INVALID;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0)
        return fail(std::string("reference scene: cannot open ") + path);
    struct stat st {};
    if (::fstat(fd, &st) != 0 || st.st_size <= 0) {
        ::close(fd);
        return fail(std::string("reference scene: empty or unreadable file ") + path);
    }
    const size_t n = (size_t)st.st_size;
    void* map = ::mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED)
        return fail(std::string("reference scene: cannot map ") + path);
    ByteReader r { (const uint8_t*)map, n };
    double info[6] = {}, cam[12] = {};
    const uint8_t* di = r.take(sizeof info); // DAGInfo (dag_info.h:6-10)
    int rc;
    if (!di) {
        rc = fail("reference scene: truncated DAGInfo");
    } else {
        std::memcpy(info, di, sizeof info);
        std::lock_guard<std::mutex> lock(ctx->mu);
        rc = read_dag(ctx, r);
    }
    if (rc == SVDAG_OK) {
        const uint8_t* c = r.take(sizeof cam);
        uint64_t numMaterials = 0;
        if (!c || !skip_textures(r, numMaterials) || r.off != n)
            rc = fail("reference scene: the camera view or voxel textures after the DAG do not parse to the end of the file");
        else
            std::memcpy(cam, c, sizeof cam);
    }
    ::munmap(map, n);
    if (rc != SVDAG_OK)
        return rc;
    if (bounds)
        std::memcpy(bounds, info, sizeof info);
    if (view)
        std::memcpy(view, cam, sizeof cam);
    return SVDAG_OK;
}

int svdag_write_reference_scene(svdag_ctx* ctx, const char* path, const double* bounds, const double* view)
{
    if (!ctx || !path)
        return SVDAG_E_INVALID;
    const double side = (double)(1ull << ctx->dag.levels);
    const double defInfo[6] = { 0.0, 0.0, 0.0, side, side, side };
    const double defView[12] = { 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1 }; // CameraView defaults (camera_view.h:12-13)
    ByteWriter w;
    const double* info = bounds ? bounds : defInfo;
    const double* cam = view ? view : defView;
    for (int i = 0; i < 6; ++i)
        w.put<double>(info[i]);
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        SVDAG_TRY(write_dag(ctx, w));
    }
    for (int i = 0; i < 12; ++i)
        w.put<double>(cam[i]);
    w.put<uint64_t>(0); // no material names
    w.put<uint64_t>(0); // no material textures
    FILE* f = std::fopen(path, "wb");
    if (!f)
        return fail(std::string("reference scene: cannot create ") + path);
    bool ok = std::fwrite(w.b.data(), 1, w.b.size(), f) == w.b.size();
    ok &= std::fclose(f) == 0;
    if (!ok)
        return fail(std::string("reference scene: write failed: ") + path);
    return SVDAG_OK;
}

} // extern "C"
