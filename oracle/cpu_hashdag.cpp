// The reference's CPU HashDAG editing path, geometry only, restated in C++ as
// the CPU timing baseline of the GPU back-end.  Test infrastructure: only
// tests/ and bench.py's cpu_baseline legs load it (through liboracle.so).
//
// It follows, in mathijs727/GPU-SVDAG-Editing:
//  - HashTable (src/dags/hash_dag/hash_table.h): one region of buckets per
//    level, each bucket a fixed span of virtual words (levels 0..8: 1024
//    buckets of 1024 words; deeper levels: 65536 buckets of 4096 words;
//    typedefs.h:229-265, hash_dag_globals.h), virtual pages of 512 words
//    mapped on first use through a page table into a pool (allocate_page,
//    hash_table.h:748-765), items appended to their bucket and found by a
//    linear scan of it (find_leaf_node_in_bucket 188-247,
//    find_interior_node_in_bucket 249-334), find_or_add_* with a per-bucket
//    mutex and a second find when the bucket grew meanwhile (443-528), and
//    interior nodes padded to the next page when they would straddle one
//    (add_interior_node 383-441).
//  - Hashes (hash_table.h:71-80; utils.h:351-383 with USE_ALTERNATE_HASH=1):
//    murmurhash64 of the leaf mask, murmurhash32xN of the interior node.
//  - add_full_node (hash_table.h:708-745): the full leaf and one full node
//    per level go in first (ADD_FULL_NODES_FIRST).
//  - HashDagEdits::edit_threaded / edit_leaf (hash_dag_edits.h:106-643) and
//    HashDAG::edit_threads (hash_dag.h:237-378): a first pass queues one task
//    per edited child at level C_colorTreeLevels, NUM_THREADS std::threads
//    (created per edit) take tasks from an atomic counter and run edit_leaf,
//    and a second pass assembles the levels above.
//  - SphereEditor<bool>, BoxEditor<bool> (hash_dag_editors.h:56-119, 211-280).
// Configuration: the reference defaults with EDITS_ENABLE_COLORS=0 and
// UNDO_REDO=0 (its benchmark settings, python/shared.py:24-46), no Bloom
// filter (USE_BLOOM_FILTER's default 0, typedefs.h:336), EARLY_FULL_CHECK on.
// Declared deviations: C_colorTreeLevels = min(10, levels - 3) (the
// reference's fixed 10 fails its own voxel-count bound below depth 13), and
// bucket sizes are parameters (the reference's are compile-time knobs) that
// the caller may raise when a scene would overflow the default spans; an
// overflowing bucket is reported as an error instead of corrupting memory.
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint32_t kInvalid = 0xDFFFFFFu;   // C_InvalidNodeIndex (hash_dag_edits.h:17)
constexpr uint32_t kNotFound = 0xFFFFFFFFu;
constexpr uint32_t kPageSize = 512;         // PAGE_SIZE (typedefs.h:229)

inline uint32_t popc(uint32_t v) { return (uint32_t)__builtin_popcount(v); }
inline uint32_t popcll(uint64_t v) { return (uint32_t)__builtin_popcountll(v); }
inline uint32_t total_size(uint32_t header) { return popc(header & 0xFFu) + 1u; } // utils.h:334-337

// utils.h:351-359
inline uint64_t murmurhash64(uint64_t h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}
// utils.h:362-383 (USE_ALTERNATE_HASH)
inline uint32_t murmurhash32xN(const uint32_t* p, uint32_t n)
{
    uint32_t h = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k = p[i];
        k *= 0xcc9e2d51u;
        k = (k << 15) | (k >> 17);
        k *= 0x1b873593u;
        h ^= k;
        h = (h << 13) | (h >> 19);
        h = h * 5u + 0xe6546b64u;
    }
    h ^= n;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

struct Path {
    uint32_t x = 0, y = 0, z = 0;
    void descend(uint32_t c) // path.h:27-35
    {
        x = (x << 1) | ((c & 4u) >> 2);
        y = (y << 1) | ((c & 2u) >> 1);
        z = (z << 1) | (c & 1u);
    }
    uint64_t key() const { return ((uint64_t)x << 42) ^ ((uint64_t)y << 21) ^ z; }
};

// SphereEditor<isAdding> / BoxEditor<isAdding> over boxes [start, start + 2^depth].
struct Editor {
    bool sphere = true, adding = true;
    float cx, cy, cz, radius, r2;     // sphere
    float bmin[3], bmax[3];           // box
    bool should_edit_impl(const float* mn, const float* mx) const
    {
        if (sphere) { // hash_dag_editors.h:97-118
            float dist = r2;
            const float c[3] = { cx, cy, cz };
            for (int a = 0; a < 3; ++a) {
                if (c[a] < mn[a])
                    dist -= (c[a] - mn[a]) * (c[a] - mn[a]);
                else if (c[a] > mx[a])
                    dist -= (c[a] - mx[a]) * (c[a] - mx[a]);
            }
            return dist > 0;
        }
        for (int a = 0; a < 3; ++a) // hash_dag_editors.h:74-78
            if (bmin[a] >= mx[a] || bmax[a] <= mn[a] || mn[a] >= bmax[a] || mx[a] <= bmin[a])
                return false;
        return true;
    }
    bool contained(const float* mn, const float* mx) const
    {
        if (sphere) { // hash_dag_editors.h:256-263
            const float c[3] = { cx, cy, cz };
            float s = 0.0f;
            for (int a = 0; a < 3; ++a)
                s += std::max((mn[a] - c[a]) * (mn[a] - c[a]), (mx[a] - c[a]) * (mx[a] - c[a]));
            return s < r2;
        }
        for (int a = 0; a < 3; ++a) // hash_dag_editors.h:220-227
            if (!(bmin[a] <= mn[a] && bmax[a] >= mx[a]))
                return false;
        return true;
    }
    static void box_of(const Path& p, uint32_t depth, float* mn, float* mx)
    {
        const uint32_t v[3] = { p.x, p.y, p.z };
        for (int a = 0; a < 3; ++a) {
            mn[a] = (float)(v[a] << depth);
            mx[a] = mn[a] + (float)(1u << depth);
        }
    }
    bool should_edit(const Path& p, uint32_t depth) const
    {
        float mn[3], mx[3];
        box_of(p, depth, mn, mx);
        return should_edit_impl(mn, mx);
    }
    bool is_full(const Path& p, uint32_t depth) const
    {
        if (!adding)
            return false;
        float mn[3], mx[3];
        box_of(p, depth, mn, mx);
        return contained(mn, mx);
    }
    bool is_empty(const Path& p, uint32_t depth) const
    {
        if (adding)
            return false;
        float mn[3], mx[3];
        box_of(p, depth, mn, mx);
        return contained(mn, mx);
    }
};

struct HashDag {
    uint32_t levels = 0, leafLevel = 0, colorLevels = 0;
    uint32_t numTop = 9, bucketsTop = 1024, bucketsLow = 65536, sizeTop = 1024, sizeLow = 4096;
    std::vector<uint64_t> levelBase;     // first virtual word of each level
    std::vector<uint32_t> levelBucket0;  // first global bucket index of each level
    uint64_t virtualWords = 0;
    std::unique_ptr<std::atomic<uint32_t>[]> bucketSizes;
    std::unique_ptr<std::mutex[]> mutexes;
    std::mutex pageMutex;
    std::unique_ptr<std::atomic<uint32_t>[]> pageTable; // virtual page -> pool page (0 = unmapped)
    uint64_t numPages = 0;
    uint32_t* pool = nullptr;
    uint64_t poolMaxPages = 0;
    uint32_t poolTop = 1;
    std::vector<uint32_t> fullNode;
    uint32_t root = kInvalid;
    std::atomic<uint32_t> overflow { 0 };

    ~HashDag()
    {
        if (pool)
            munmap(pool, poolMaxPages * kPageSize * 4);
    }

    uint32_t buckets(uint32_t level) const { return level < numTop ? bucketsTop : bucketsLow; }
    uint32_t bucket_span(uint32_t level) const { return level < numTop ? sizeTop : sizeLow; }

    bool init(uint32_t L, uint32_t colorL, uint32_t topSpan, uint32_t lowSpan, uint64_t poolBytes)
    {
        levels = L;
        leafLevel = L - 2;
        colorLevels = colorL;
        if (topSpan)
            sizeTop = topSpan;
        if (lowSpan)
            sizeLow = lowSpan;
        if (sizeTop % kPageSize || sizeLow % kPageSize)
            return false;
        uint64_t base = 0, b0 = 0;
        for (uint32_t l = 0; l < levels; ++l) {
            levelBase.push_back(base);
            levelBucket0.push_back((uint32_t)b0);
            base += (uint64_t)buckets(l) * bucket_span(l);
            b0 += buckets(l);
        }
        virtualWords = base;
        if (virtualWords >= (1ull << 32))
            return false; // pointers are 32-bit virtual word addresses
        bucketSizes.reset(new std::atomic<uint32_t>[b0]);
        for (uint64_t i = 0; i < b0; ++i)
            bucketSizes[i].store(0, std::memory_order_relaxed);
        mutexes.reset(new std::mutex[b0]);
        numPages = virtualWords / kPageSize;
        pageTable.reset(new std::atomic<uint32_t>[numPages]);
        for (uint64_t i = 0; i < numPages; ++i)
            pageTable[i].store(0, std::memory_order_relaxed);
        poolMaxPages = poolBytes / (kPageSize * 4);
        void* p = mmap(nullptr, poolMaxPages * kPageSize * 4, PROT_READ | PROT_WRITE,
                       MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (p == MAP_FAILED)
            return false;
        pool = static_cast<uint32_t*>(p);
        fullNode.assign(levels, kInvalid);
        for (uint32_t l = leafLevel + 1; l-- > 0;)
            add_full_node(l);
        return overflow.load() == 0;
    }

    uint32_t* sys_ptr(uint32_t ptr) const // get_sys_ptr_impl (hash_table.h:146-163)
    {
        const uint32_t page = ptr / kPageSize;
        return pool + (uint64_t)pageTable[page].load(std::memory_order_relaxed) * kPageSize + ptr % kPageSize;
    }
    uint32_t make_ptr(uint32_t level, uint32_t bucket, uint32_t pos) const
    {
        return (uint32_t)(levelBase[level] + (uint64_t)bucket * bucket_span(level) + pos);
    }
    std::atomic<uint32_t>& bucket_size(uint32_t level, uint32_t bucket) const
    {
        return bucketSizes[levelBucket0[level] + bucket];
    }
    void allocate_page(uint32_t page) // hash_table.h:748-765
    {
        std::lock_guard<std::mutex> lock(pageMutex);
        if (pageTable[page].load(std::memory_order_relaxed) != 0)
            return;
        if (poolTop >= poolMaxPages) {
            overflow.store(1);
            return;
        }
        pageTable[page].store(poolTop++, std::memory_order_release);
    }

    uint32_t find_leaf_in_bucket(uint32_t bucketSize, uint32_t level, uint32_t bucket, uint64_t bits) const
    {
        const uint32_t base = make_ptr(level, bucket, 0);
        for (uint32_t p = 0; p < bucketSize; p += kPageSize) {
            const uint32_t* sp = sys_ptr(base + p);
            const uint32_t end = std::min(bucketSize, p + kPageSize);
            for (uint32_t i = p; i < end; i += 2, sp += 2) {
                uint64_t v;
                std::memcpy(&v, sp, 8);
                if (v == bits)
                    return base + i;
            }
        }
        return kNotFound;
    }
    // The reference's page bound: a page whose used part is not longer than
    // the needle ends the search, and the scan stops nodeSize words before
    // the page's used end (hash_table.h:301-321), as written there.
    uint32_t find_interior_in_bucket(uint32_t bucketSize, uint32_t level, uint32_t bucket, uint32_t n,
                                     const uint32_t* node) const
    {
        const uint32_t base = make_ptr(level, bucket, 0);
        for (uint32_t p = 0; p < bucketSize; p += kPageSize) {
            const uint32_t* sp = sys_ptr(base + p);
            uint32_t pageEnd = std::min(bucketSize, p + kPageSize);
            if (p + n >= pageEnd)
                return kNotFound;
            pageEnd -= n;
            for (uint32_t i = p; i < pageEnd;) {
                if (std::memcmp(node, sp, n * 4) == 0)
                    return base + i;
                const uint32_t len = total_size(*sp);
                i += len;
                sp += len;
            }
        }
        return kNotFound;
    }
    uint32_t add_leaf(uint32_t level, uint64_t leaf, uint32_t hash) // hash_table.h:336-368
    {
        const uint32_t bucket = hash & (buckets(level) - 1u);
        std::atomic<uint32_t>& bs = bucket_size(level, bucket);
        const uint32_t size = bs.load(std::memory_order_relaxed);
        if (size + 2 > bucket_span(level)) {
            overflow.store(1);
            return kInvalid;
        }
        const uint32_t ptr = make_ptr(level, bucket, size);
        const uint32_t page = ptr / kPageSize;
        if (size % kPageSize == 0 && pageTable[page].load(std::memory_order_acquire) == 0)
            allocate_page(page);
        if (overflow.load(std::memory_order_relaxed))
            return kInvalid;
        std::memcpy(sys_ptr(ptr), &leaf, 8);
        bs.store(size + 2, std::memory_order_release);
        return ptr;
    }
    uint32_t add_interior(uint32_t level, uint32_t n, const uint32_t* node, uint32_t hash) // 369-441
    {
        const uint32_t bucket = hash & (buckets(level) - 1u);
        std::atomic<uint32_t>& bs = bucket_size(level, bucket);
        uint32_t size = bs.load(std::memory_order_relaxed);
        const uint32_t left = kPageSize - size % kPageSize;
        if (left == kPageSize || left < n) {
            if (left != kPageSize)
                size += left;
            if (size + n > bucket_span(level)) {
                overflow.store(1);
                return kInvalid;
            }
            const uint32_t page = make_ptr(level, bucket, size) / kPageSize;
            if (pageTable[page].load(std::memory_order_acquire) == 0)
                allocate_page(page);
            if (overflow.load(std::memory_order_relaxed))
                return kInvalid;
        }
        if (size + n > bucket_span(level)) {
            overflow.store(1);
            return kInvalid;
        }
        const uint32_t ptr = make_ptr(level, bucket, size);
        std::memcpy(sys_ptr(ptr), node, n * 4);
        bs.store(size + n, std::memory_order_release);
        return ptr;
    }
    uint32_t find_or_add_leaf(uint64_t leaf) // hash_table.h:443-480
    {
        const uint32_t level = leafLevel;
        const uint32_t hash = (uint32_t)murmurhash64(leaf);
        const uint32_t bucket = hash & (buckets(level) - 1u);
        const uint32_t size = bucket_size(level, bucket).load(std::memory_order_acquire);
        uint32_t r = find_leaf_in_bucket(size, level, bucket, leaf);
        if (r == kNotFound) {
            std::lock_guard<std::mutex> lock(mutexes[levelBucket0[level] + bucket]);
            const uint32_t now = bucket_size(level, bucket).load(std::memory_order_acquire);
            if (now != size)
                r = find_leaf_in_bucket(now, level, bucket, leaf);
            if (r == kNotFound)
                r = add_leaf(level, leaf, hash);
        }
        return r;
    }
    uint32_t find_or_add_interior(uint32_t level, uint32_t n, const uint32_t* node) // hash_table.h:481-528
    {
        const uint32_t hash = murmurhash32xN(node, n);
        const uint32_t bucket = hash & (buckets(level) - 1u);
        const uint32_t size = bucket_size(level, bucket).load(std::memory_order_acquire);
        uint32_t r = find_interior_in_bucket(size, level, bucket, n, node);
        if (r == kNotFound) {
            std::lock_guard<std::mutex> lock(mutexes[levelBucket0[level] + bucket]);
            const uint32_t now = bucket_size(level, bucket).load(std::memory_order_acquire);
            if (now != size)
                r = find_interior_in_bucket(now, level, bucket, n, node);
            if (r == kNotFound)
                r = add_interior(level, n, node, hash);
        }
        return r;
    }
    void add_full_node(uint32_t level) // hash_table.h:708-745
    {
        if (level == leafLevel) {
            const uint64_t leaf = ~0ull;
            fullNode[level] = add_leaf(level, leaf, (uint32_t)murmurhash64(leaf));
            return;
        }
        const uint32_t count = level < colorLevels ? 0u : (1u << (3 * (levels - level)));
        uint32_t node[9];
        node[0] = (count << 8) | 0xFFu;
        for (uint32_t c = 0; c < 8; ++c)
            node[1 + c] = fullNode[level + 1];
        fullNode[level] = add_interior(level, 9, node, murmurhash32xN(node, 9));
    }

    uint32_t voxel_count(uint32_t level, uint32_t index) const // edit_leaf's getVoxelCount (296-308)
    {
        const uint32_t* p = sys_ptr(index);
        return level < leafLevel ? p[0] >> 8 : popc(p[0]) + popc(p[1]);
    }
};

struct LeafResult {
    uint32_t index, count;
};

// HashDagEdits::edit_leaf (hash_dag_edits.h:283-642), geometry only.
LeafResult edit_leaf(HashDag& d, uint32_t level, uint32_t nodeIndex, Path path, const Editor& ed)
{
    const uint32_t depth = d.levels - level;
    if (!ed.should_edit(path, depth))
        return { nodeIndex, nodeIndex != kInvalid ? d.voxel_count(level, nodeIndex) : 0u };
    if (ed.is_empty(path, depth))
        return { kInvalid, 0 };
    if (ed.is_full(path, depth))
        return { d.fullNode[level], 1u << (3 * depth) };
    if (level == d.leafLevel) {
        uint64_t mask = 0;
        if (nodeIndex != kInvalid)
            std::memcpy(&mask, d.sys_ptr(nodeIndex), 8);
        const uint64_t initial = mask;
        for (uint32_t c1 = 0; c1 < 8; ++c1) {
            Path p1 = path;
            p1.descend(c1);
            for (uint32_t c2 = 0; c2 < 8; ++c2) {
                Path p2 = p1;
                p2.descend(c2);
                float mn[3], mx[3];
                Editor::box_of(p2, 0, mn, mx);
                const uint64_t bit = 1ull << (c1 * 8u + c2);
                if (ed.should_edit_impl(mn, mx)) {
                    if (ed.adding)
                        mask |= bit;
                    else
                        mask &= ~bit;
                }
            }
        }
        if (mask == initial)
            return { nodeIndex, popcll(mask) };
        if (mask == 0)
            return { kInvalid, 0 };
        return { d.find_or_add_leaf(mask), popcll(mask) };
    }
    const uint32_t* np = nodeIndex != kInvalid ? d.sys_ptr(nodeIndex) : nullptr;
    const uint32_t childMask = np ? (np[0] & 0xFFu) : 0u;
    uint32_t off = 1, count = 0, children[8];
    bool changed = false;
    for (uint32_t c = 0; c < 8; ++c) {
        Path p = path;
        p.descend(c);
        const uint32_t old = (childMask >> c) & 1u ? np[off++] : kInvalid;
        const LeafResult r = edit_leaf(d, level + 1, old, p, ed);
        if (r.index != kInvalid)
            count += r.count;
        changed |= old != r.index;
        children[c] = r.index;
    }
    if (!changed)
        return { nodeIndex, count };
    uint32_t node[9], n = 0;
    node[n++] = count << 8;
    for (uint32_t c = 0; c < 8; ++c)
        if (children[c] != kInvalid) {
            node[n++] = children[c];
            node[0] |= 1u << c;
        }
    if ((node[0] & 0xFFu) == 0)
        return { kInvalid, 0 };
    return { d.find_or_add_interior(level, n, node), count };
}

struct Task {
    uint32_t level, nodeIndex;
    Path path;
    uint32_t result = kInvalid;
};

struct EditParams {
    std::vector<Task> tasks;
    std::unordered_map<uint64_t, uint32_t> byPath; // path -> task (the reference's threads map)
};

// HashDagEdits::edit_threaded (hash_dag_edits.h:106-281) without colours.
template <bool Queue>
uint32_t edit_threaded(HashDag& d, uint32_t level, uint32_t nodeIndex, Path path, const Editor& ed, EditParams& P)
{
    if (!ed.should_edit(path, d.levels - level))
        return Queue ? kInvalid : nodeIndex;
    const uint32_t* np = nodeIndex != kInvalid ? d.sys_ptr(nodeIndex) : nullptr;
    const uint32_t childMask = np ? (np[0] & 0xFFu) : 0u;
    uint32_t off = 1, children[8];
    bool changed = false;
    for (uint32_t c = 0; c < 8; ++c) {
        Path p = path;
        p.descend(c);
        const uint32_t old = (childMask >> c) & 1u ? np[off++] : kInvalid;
        uint32_t now;
        if (level == d.colorLevels - 1) {
            if (ed.should_edit(p, d.levels - level - 1)) {
                if (Queue) {
                    P.byPath.emplace(p.key(), (uint32_t)P.tasks.size());
                    P.tasks.push_back({ level + 1, old, p });
                    now = kInvalid;
                } else {
                    now = P.tasks[P.byPath.at(p.key())].result;
                }
            } else {
                now = Queue ? kInvalid : old;
            }
        } else {
            now = edit_threaded<Queue>(d, level + 1, old, p, ed, P);
        }
        if (!Queue) {
            changed |= old != now;
            children[c] = now;
        }
    }
    if (Queue)
        return kInvalid;
    if (!changed)
        return nodeIndex;
    uint32_t node[9], n = 0;
    node[n++] = 0; // no voxel counts above the colour-tree levels
    for (uint32_t c = 0; c < 8; ++c)
        if (children[c] != kInvalid) {
            node[n++] = children[c];
            node[0] |= 1u << c;
        }
    if ((node[0] & 0xFFu) == 0)
        return kInvalid;
    return d.find_or_add_interior(level, n, node);
}

// HashDAG::edit_threads (hash_dag.h:237-378): first pass, NUM_THREADS
// workers over the queued tasks, second pass.
void edit_threads(HashDag& d, const Editor& ed, uint32_t numThreads, uint64_t* tasksOut)
{
    EditParams P;
    edit_threaded<true>(d, 0, d.root, Path {}, ed, P);
    std::atomic<uint32_t> next { 0 };
    auto worker = [&]() {
        for (;;) {
            const uint32_t t = next.fetch_add(1);
            if (t >= P.tasks.size())
                return;
            Task& k = P.tasks[t];
            k.result = edit_leaf(d, k.level, k.nodeIndex, k.path, ed).index;
        }
    };
    if (numThreads <= 1) {
        worker();
    } else {
        std::vector<std::thread> threads;
        threads.reserve(numThreads);
        for (uint32_t i = 0; i < numThreads; ++i)
            threads.emplace_back(worker);
        for (auto& t : threads)
            t.join();
    }
    d.root = edit_threaded<false>(d, 0, d.root, Path {}, ed, P);
    if (tasksOut)
        *tasksOut = P.tasks.size();
}

} // namespace

extern "C" {

// levels: D; color_levels 0 = min(10, D - 3); top_span / low_span: words per
// bucket of levels 0..8 / deeper (0 = 1024 / 4096); pool_bytes: pool reserve.
void* hd_create(uint32_t levels, uint32_t color_levels, uint32_t top_span, uint32_t low_span, uint64_t pool_bytes)
{
    if (levels < 4 || levels > 20)
        return nullptr;
    auto* d = new HashDag();
    const uint32_t cl = color_levels ? color_levels : std::min(10u, levels - 3u);
    if (cl < 1 || cl > levels - 2 || levels - cl > 7 || !d->init(levels, cl, top_span, low_span, pool_bytes)) {
        delete d;
        return nullptr;
    }
    return d;
}

void hd_destroy(void* h) { delete static_cast<HashDag*>(h); }

// Geometry of a canonical dump (for level = leaf level down to 0: [level,
// count, then per item size, words...], then the root id), imported bottom up
// with find_or_add on `threads` threads per level.  Leaves keep words 0-1
// (the mask); nodes keep the child mask, their children mapped, and the
// voxel count in the header from the colour-tree levels down.  0 = ok.
int hd_import_canonical(void* h, const uint32_t* dump, uint64_t len, uint32_t threads)
{
    auto* d = static_cast<HashDag*>(h);
    uint64_t p = 0;
    std::vector<uint32_t> below, belowCount;
    for (int L = (int)d->leafLevel; L >= 0; --L) {
        if (p + 2 > len || dump[p] != (uint32_t)L)
            return -1;
        const uint32_t n = dump[p + 1];
        p += 2;
        std::vector<uint64_t> at(n);
        for (uint32_t i = 0; i < n; ++i) {
            if (p >= len || p + 1 + dump[p] > len)
                return -1;
            at[i] = p;
            p += 1 + dump[p];
        }
        std::vector<uint32_t> ids(n), counts(n);
        std::atomic<uint32_t> next { 0 }, bad { 0 };
        auto work = [&]() {
            for (;;) {
                const uint32_t i = next.fetch_add(1);
                if (i >= n)
                    return;
                const uint32_t* w = dump + at[i] + 1;
                const uint32_t s = dump[at[i]];
                if ((uint32_t)L == d->leafLevel) {
                    const uint64_t mask = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
                    ids[i] = d->find_or_add_leaf(mask);
                    counts[i] = popcll(mask);
                    continue;
                }
                uint32_t node[9], cnt = 0;
                const uint32_t m = w[0] & 0xFFu;
                if (s != popc(m) + 1) {
                    bad.store(1);
                    return;
                }
                for (uint32_t j = 1; j < s; ++j) {
                    if (w[j] >= below.size()) {
                        bad.store(1);
                        return;
                    }
                    node[j] = below[w[j]];
                    cnt += belowCount[w[j]];
                }
                node[0] = m | ((uint32_t)L >= d->colorLevels ? cnt << 8 : 0u);
                ids[i] = d->find_or_add_interior((uint32_t)L, s, node);
                counts[i] = cnt;
            }
        };
        std::vector<std::thread> pool;
        for (uint32_t t = 1; t < std::max(1u, threads); ++t)
            pool.emplace_back(work);
        work();
        for (auto& t : pool)
            t.join();
        if (bad.load() || d->overflow.load())
            return -2;
        below.swap(ids);
        belowCount.swap(counts);
    }
    if (p + 1 != len)
        return -1;
    d->root = dump[p] == 0xFFFFFFFFu ? kInvalid : below.at(dump[p]);
    return 0;
}

// One edit: kind 0 sphere add, 1 sphere remove, 2 box add, 3 box remove
// (centre, radius as the reference tools take them).  *ms = wall time of the
// edit, *tasks = leaf tasks run.  0 = ok, -2 = a bucket or the pool overflowed.
int hd_edit(void* h, uint32_t kind, float cx, float cy, float cz, float r, uint32_t threads, double* ms,
            uint64_t* tasks)
{
    auto* d = static_cast<HashDag*>(h);
    Editor e;
    e.sphere = kind < 2;
    e.adding = kind == 0 || kind == 2;
    e.cx = cx;
    e.cy = cy;
    e.cz = cz;
    e.radius = r;
    e.r2 = r * r;
    const float c[3] = { cx, cy, cz };
    for (int a = 0; a < 3; ++a) {
        e.bmin[a] = c[a] - r;
        e.bmax[a] = c[a] + r;
    }
    const auto t0 = std::chrono::steady_clock::now();
    edit_threads(*d, e, threads, tasks);
    const auto t1 = std::chrono::steady_clock::now();
    if (ms)
        *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return d->overflow.load() ? -2 : 0;
}

// 1 where the voxel is set (DAGUtils::get_value over the CPU HashDAG).
void hd_query(void* h, const uint32_t* xyz, uint64_t n, uint8_t* out)
{
    auto* d = static_cast<HashDag*>(h);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        uint32_t idx = d->root;
        uint8_t v = 0;
        for (uint32_t level = 0; idx != kInvalid; ++level) {
            const uint32_t* p = d->sys_ptr(idx);
            if (level == d->leafLevel) {
                const uint32_t bit = (((x >> 1) & 1u) << 5) | (((y >> 1) & 1u) << 4) | (((z >> 1) & 1u) << 3)
                    | ((x & 1u) << 2) | ((y & 1u) << 1) | (z & 1u);
                const uint64_t mask = (uint64_t)p[0] | ((uint64_t)p[1] << 32);
                v = (uint8_t)((mask >> bit) & 1u);
                break;
            }
            const uint32_t sh = d->levels - 1 - level;
            const uint32_t c = (((x >> sh) & 1u) << 2) | (((y >> sh) & 1u) << 1) | ((z >> sh) & 1u);
            const uint32_t m = p[0] & 0xFFu;
            if (!((m >> c) & 1u))
                break;
            idx = p[1 + popc(m & ((1u << c) - 1u))];
        }
        out[i] = v;
    }
}

// Words stored in the tables (bucket fill, page padding included) and pool pages mapped.
void hd_usage(void* h, uint64_t out[2])
{
    auto* d = static_cast<HashDag*>(h);
    uint64_t words = 0;
    for (uint32_t l = 0; l < d->levels; ++l)
        for (uint32_t b = 0; b < d->buckets(l); ++b)
            words += d->bucket_size(l, b).load();
    out[0] = words;
    out[1] = d->poolTop - 1u;
}

} // extern "C"
