// Device-side primitives shared by every kernel of the SVDAG editing back-end:
// item encoding, the wave64 slab hash table (find / insert, group-cooperative
// and per-lane forms), and the slab allocator.
//
// Encoding follows GPUBaseDAG (dags/my_gpu_dags/my_gpu_base_dag.h:86-351):
//   node header = mask[0:8] | fullFlag[8] | fullMaterial[9:9+M] | payload[9+M:32]
//   node        = header + one handle per set mask bit (2..9 words)
//   leaf        = [mask lo][mask hi][packed M-bit materials in voxel-bit order]
//   handle      = index[0:32-b] | (size - minSize)[32-b:32]
// Tables are indexed by item size - 2, nodes and leaves of equal word count
// share one table (my_gpu_hash_dag.h:173, 240).
//
// Table layout (own design, not the reference's 32-slot slab): one table per
// item size s; a slab is one 128-byte line holding everything a probe needs:
//   words [0, tagWords)                        one-byte tags (0 = empty), one per slot
//   word  tagWords                             next slab id of the bucket's chain
//   words [tagWords + 1, tagWords + 1 + slots*s) the items themselves
// where slots is the most items of s words that fit beside their tags and
// the link (13 at s = 2, 5 at s = 5, 3 at s = 9, 1 at s >= 16).  The tags and
// the link sit in the line's first 32 bytes, so a probe that matches no tag
// (every miss in a chainless bucket) reads one 32-byte sector, and the
// compare of a candidate reads the same line.  Head slabs 0..numBuckets-1 are the buckets; overflow slabs come
// from the slab allocator.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svdag {

constexpr uint32_t kLineWords = 32;     // words per line (128 B)
constexpr uint32_t kItemBudget = 31;    // words for tags and items (the link takes one)
constexpr uint32_t kSlotBits = 4;       // pointer = slab << 4 | slot
constexpr uint32_t kMaxSlots = 1u << kSlotBits; // bound on slots per line (13 at most)
constexpr uint32_t kNoSlab = 0xFFFFFFFFu;
constexpr uint32_t kPendingSlab = 0xFFFFFFFEu; // tail link while its successor is being allocated
constexpr uint32_t kNotFound = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFFu;
constexpr uint32_t kMaxItemWords = 18;  // 2 + 64*8/32 (8 material bits)
constexpr uint8_t kRoleLeafUse = 0x80;   // find filter: an item without roles (never stored)

// Items of s words per line: the largest k with k*s + ceil(k/4) <= 31.
__host__ __device__ inline uint32_t line_slots(uint32_t s)
{
    uint32_t k = 1;
    while ((k + 1) * s + (k + 4) / 4 <= kItemBudget)
        ++k;
    return k;
}
__host__ __device__ inline uint32_t line_tag_words(uint32_t slots)
{
    return (slots + 3) / 4;
}

// Bucket hash: SlabHash-style xor of per-word modular hashes (utils.h:386-403).
constexpr uint32_t kHashX = 2717583853u;
constexpr uint32_t kHashY = 1117415409u;
constexpr uint32_t kPrime = 4294967291u;

__host__ __device__ inline uint32_t hash_part(uint32_t w, uint32_t j)
{
    const uint32_t x = kHashX * w + j * kHashY; // u32 wrap, then mod prime
    return x >= kPrime ? x - kPrime : x;
}
// One-byte tag stored per slot; 0 marks an empty slot.
__host__ __device__ inline uint32_t make_tag(uint32_t h)
{
    const uint32_t t = (h * 0x9E3779B1u) >> 24;
    return t ? t : 1u;
}
__host__ __device__ inline uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
// Exact per-byte equality-to-zero mask: bit 7 of every zero byte.
__host__ __device__ inline uint32_t zero_bytes(uint32_t x)
{
    const uint32_t y = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return ~(y | x | 0x7F7F7F7Fu);
}
// Compact bits 7/15/23/31 into bits 0..3.
__host__ __device__ inline uint32_t compact_bytes(uint32_t z)
{
    return ((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u);
}

// ---------------------------------------------------------------- allocator
// Slab ids come from a free stack (ids in freeList[0, freeCount)) and, when
// it is empty, from the never-used tail of the pre-reserved arena (bump
// counter `fresh`).  Pops (chain growth during inserts) and pushes (the GC
// sweep) never run in the same kernel, so both are plain counter moves,
// aggregated to one atomic per wave.  A pop may drive freeCount below zero
// (the lanes past the stack's end take fresh ids instead); the GC clamps it
// back to zero before it pushes.
struct AllocState {
    uint32_t fresh;     // next never-used slab id (starts at numBuckets)
    int32_t freeCount;  // ids on the free stack; <= 0 when empty
    uint32_t oom;       // set when an allocation fails
    uint32_t pad;
};

// Every line that is not on a chain (the arena's never-used tail and the
// free stack) is empty: tag words 0 and next = kNoSlab.
struct TableDev {
    uint32_t* lines;     // [slabCap][32]: tags, next, items
    uint32_t* freeList;  // [slabCap] free stack of slab ids
    AllocState* alloc;
    uint32_t* refs;      // DAG tables: [slabCap][16] stored references to each slot's item
    uint8_t* roles;      // DAG tables: [slabCap][16] kRoleParentOf* bits of each slot's item
    uint32_t itemWords;
    uint32_t numBuckets;
    uint32_t slabCap;
    uint32_t slots;      // items per line
    uint32_t tagWords;   // tag words at the line start
    uint32_t pad[3];
};

__device__ inline uint32_t ld_relaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_relaxed64(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Loads that may be served by the CU's caches and so may return an older
// value.  Only for words that change monotonically within a kernel (bits
// set, a slot going from empty to claimed), as a filter in front of the
// agent-scope atomic that decides: a stale value only sends the caller to
// that atomic.
__device__ inline uint32_t ld_filter(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline unsigned long long ld_filter64(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Lanes below this one in the wave.
__device__ inline unsigned long long lanes_below()
{
    const uint32_t l = __lane_id();
    return l ? (~0ull >> (64u - l)) : 0ull;
}

// Slabs for the lanes of mask m, all of one table: one atomic on the
// free-stack count and at most one on the bump counter.
__device__ inline uint32_t slab_alloc_lanes(const TableDev& t, unsigned long long m)
{
    const uint32_t k = (uint32_t)__popcll(m);
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
    const uint32_t rank = (uint32_t)__popcll(m & lanes_below());
    AllocState* a = t.alloc;
    int top = 0;
    uint32_t base = 0;
    if (__lane_id() == leader) {
        top = atomicSub(&a->freeCount, (int)k);
        const uint32_t avail = top <= 0 ? 0u : ((uint32_t)top < k ? (uint32_t)top : k);
        if (avail < k)
            base = atomicAdd(&a->fresh, k - avail);
    }
    top = __shfl(top, (int)leader, 64);
    base = (uint32_t)__shfl((int)base, (int)leader, 64);
    const uint32_t avail = top <= 0 ? 0u : ((uint32_t)top < k ? (uint32_t)top : k);
    if (rank < avail)
        return ld_relaxed(&t.freeList[(uint32_t)top - 1u - rank]);
    const uint32_t id = base + (rank - avail);
    if (id >= t.slabCap) {
        atomicOr(&a->oom, 1u);
        return kNoSlab;
    }
    return id;
}

// A slab for every active lane that passes need = true, from that lane's
// table.  Lanes of one wave may insert into different tables (items of
// different sizes), so requests are served one table at a time, each with
// one aggregated allocation.  Lanes passing false get kNoSlab; kNoSlab with
// oom set when the arena is exhausted.
__device__ inline uint32_t wave_slab_alloc(const TableDev& t, bool need)
{
    unsigned long long pending = __ballot(need);
    const uint64_t mine = (uint64_t)(uintptr_t)t.alloc;
    uint32_t id = kNoSlab;
    while (pending != 0ull) {
        const int first = __ffsll((long long)pending) - 1;
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)mine, first, 64);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(mine >> 32), first, 64);
        const bool take = ((pending >> __lane_id()) & 1ull) != 0ull && mine == (((uint64_t)hi << 32) | lo);
        const unsigned long long m = __ballot(take);
        pending &= ~m;
        if (take)
            id = slab_alloc_lanes(t, m);
    }
    return id;
}

// Push the slabs of the active lanes passing give = true (one atomic per
// wave).  Every calling lane must use the same table t (the GC sweep runs
// one table per launch); only the sweep frees, after freeCount was clamped
// to >= 0.
__device__ inline void wave_slab_free(const TableDev& t, bool give, uint32_t id)
{
    const unsigned long long m = __ballot(give);
    if (!give)
        return;
    const uint32_t k = (uint32_t)__popcll(m);
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
    int top = 0;
    if (__lane_id() == leader)
        top = atomicAdd(&t.alloc->freeCount, (int)k);
    top = __shfl(top, (int)leader, 64);
    t.freeList[(uint32_t)top + (uint32_t)__popcll(m & lanes_below())] = id;
}

// A chain was seen full up to its tail line `tail`, whose link is kNoSlab.
// Each lane with try_ = true claims the link (CAS kNoSlab -> kPendingSlab);
// the claimers of a wave get their slabs from one aggregated allocation and
// publish them with one atomic store each.  The new line is appended at the
// tail and was already empty (tags 0, next kNoSlab) before this kernel, so
// a reader that follows the link sees a valid line without any fence.
// Returns the new slab, `tail` to re-read it (the claim was lost), or
// kNoSlab when the arena is exhausted (the link is restored).  Lanes with
// try_ = false only take part in the wave's allocation and get kNoSlab.
__device__ inline uint32_t chain_grow(const TableDev& t, uint32_t tail, bool try_)
{
    uint32_t* link = t.lines + (size_t)tail * kLineWords + t.tagWords;
    const bool mine = try_ && atomicCAS(link, kNoSlab, kPendingSlab) == kNoSlab;
    const uint32_t id = wave_slab_alloc(t, mine);
    if (!mine)
        return try_ ? tail : kNoSlab;
    st_relaxed(link, id);
    return id;
}

// ---------------------------------------------------------------- wave64 groups
// A "group" is G consecutive lanes of a wave64 that cooperate on one item.
template <int G>
__device__ inline uint32_t group_ballot(bool p, uint32_t gbase)
{
    if constexpr (G == 1) {
        return p ? 1u : 0u;
    } else {
        const unsigned long long b = __ballot(p);
        constexpr unsigned long long m = (G >= 32) ? 0xFFFFFFFFull : ((1ull << G) - 1ull);
        return (uint32_t)((b >> gbase) & m);
    }
}
template <int G>
__device__ inline uint32_t group_bcast(uint32_t v, uint32_t src)
{
    if constexpr (G == 1)
        return v;
    else
        return (uint32_t)__shfl((int)v, (int)src, G);
}
template <int G>
__device__ inline uint32_t group_xor(uint32_t v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1)
        v ^= (uint32_t)__shfl_xor((int)v, m, G);
    return v;
}

// Needle words held by a group: lane gl holds words gl, gl+G, gl+2G, ...
template <int G, int MW = (int)kMaxItemWords>
struct Needle {
    static constexpr int K = (MW + G - 1) / G;
    uint32_t w[K];
};

template <int G, int MW>
__device__ inline void needle_load(Needle<G, MW>& n, const uint32_t* p, uint32_t s, uint32_t gl)
{
#pragma unroll
    for (int k = 0; k < Needle<G, MW>::K; ++k) {
        const uint32_t j = gl + (uint32_t)k * G;
        n.w[k] = (j < s) ? p[j] : 0u;
    }
}
template <int G, int MW>
__device__ inline uint32_t needle_hash(const Needle<G, MW>& n, uint32_t s, uint32_t gl)
{
    uint32_t h = 0;
#pragma unroll
    for (int k = 0; k < Needle<G, MW>::K; ++k) {
        const uint32_t j = gl + (uint32_t)k * G;
        if (j < s)
            h ^= hash_part(n.w[k], j);
    }
    return group_xor<G>(h);
}
template <int G, int MW>
__device__ inline bool needle_equal_lane(const Needle<G, MW>& n, const uint32_t* item, uint32_t s, uint32_t gl)
{
    bool eq = true;
#pragma unroll
    for (int k = 0; k < Needle<G, MW>::K; ++k) {
        const uint32_t j = gl + (uint32_t)k * G;
        if (j < s)
            eq &= (item[j] == n.w[k]);
    }
    return eq;
}

// The tag words of a line (at most 4) and its next slab id (word tagWords):
// a 16-byte load and a word of the same 32-byte sector (the line is
// 128-byte aligned and tagWords <= 4).
struct LineHead {
    uint32_t tag[4];
    uint32_t next;
};
__device__ inline void line_head_load(LineHead& L, const uint32_t* line, uint32_t tagWords)
{
    const uint4 v = *reinterpret_cast<const uint4*>(line);
    L.tag[0] = v.x;
    L.tag[1] = v.y;
    L.tag[2] = v.z;
    L.tag[3] = v.w;
    L.next = line[tagWords];
}
// Bit 4j+b set where tag byte b of tag word j equals tag4's byte and the
// slot exists.
__device__ inline uint32_t line_match(const TableDev& t, const LineHead& L, uint32_t tag4)
{
    uint32_t cand = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j)
        if (j < t.tagWords)
            cand |= compact_bytes(zero_bytes(L.tag[j] ^ tag4)) << (4 * j);
    return cand & ((1u << t.slots) - 1u);
}
__device__ inline uint32_t line_empty(const TableDev& t, const uint32_t* tagWordsOfLine, uint32_t& empties)
{
    empties = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j)
        if (j < t.tagWords)
            empties |= compact_bytes(zero_bytes(tagWordsOfLine[j])) << (4 * j);
    empties &= (1u << t.slots) - 1u;
    return empties;
}
__host__ __device__ inline const uint32_t* line_item(const TableDev& t, uint32_t slab, uint32_t slot)
{
    return t.lines + (size_t)slab * kLineWords + t.tagWords + 1u + slot * t.itemWords;
}
__host__ __device__ inline const uint32_t* table_item_ptr(const TableDev& t, uint32_t ptr)
{
    return line_item(t, ptr >> kSlotBits, ptr & ((1u << kSlotBits) - 1u));
}

// Walk a chain from `slab` whose line head is already in L; the group
// compares candidates one at a time (every lane holds the same line head).
template <int G, int MW>
__device__ inline uint32_t probe_chain(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t tag4,
                                       uint32_t slab, LineHead L, uint32_t gl, uint32_t gbase)
{
    while (true) {
        uint32_t cand = line_match(t, L, tag4);
        while (cand) {
            const uint32_t slot = __ffs(cand) - 1u;
            cand &= cand - 1u;
            const bool eq = needle_equal_lane<G>(n, line_item(t, slab, slot), s, gl);
            if (group_ballot<G>(!eq, gbase) == 0)
                return (slab << kSlotBits) | slot;
        }
        if (L.next >= kPendingSlab)
            return kNotFound;
        slab = L.next;
        line_head_load(L, t.lines + (size_t)slab * kLineWords, t.tagWords);
    }
}

// Group-cooperative find (G in {4, 8, 16, 32}).
template <int G, int MW>
__device__ inline uint32_t table_find_group(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t h,
                                            uint32_t gl, uint32_t gbase)
{
    static_assert(G >= 4 && G <= 32, "group size");
    const uint32_t slab = h % t.numBuckets;
    LineHead L;
    line_head_load(L, t.lines + (size_t)slab * kLineWords, t.tagWords);
    return probe_chain<G, MW>(t, n, s, make_tag(h) * 0x01010101u, slab, L, gl, gbase);
}

// Group find that also reports where an insert of the needle should start:
// the first line of the chain whose tags showed a free slot, else the tail.
// The hint stays on the bucket's chain (chains only grow at the tail, and
// lines leave a chain only in a collection), so an insert that starts there
// and walks on finds every free slot it would have found from the head
// after that line -- it only skips lines the find saw full.
// role != 0: a filter on the item's role byte (DAG tables, see
// node_has_role): kRoleLeafUse matches items without roles, any other value
// items with one of its bits.
template <int G, int MW>
__device__ inline uint32_t table_find_group_hint(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t h,
                                                 uint32_t gl, uint32_t gbase, uint32_t& hint, uint8_t role = 0)
{
    static_assert(G >= 4 && G <= 32, "group size");
    const uint32_t tag4 = make_tag(h) * 0x01010101u;
    uint32_t slab = h % t.numBuckets;
    hint = kNoSlab;
    LineHead L;
    line_head_load(L, t.lines + (size_t)slab * kLineWords, t.tagWords);
    while (true) {
        uint32_t cand = line_match(t, L, tag4);
        while (cand) {
            const uint32_t slot = __ffs(cand) - 1u;
            cand &= cand - 1u;
            const bool eq = needle_equal_lane<G>(n, line_item(t, slab, slot), s, gl);
            if (group_ballot<G>(!eq, gbase) == 0) {
                const uint32_t ptr = (slab << kSlotBits) | slot;
                if (role == 0 || (role == kRoleLeafUse ? t.roles[ptr] == 0 : (t.roles[ptr] & role) != 0))
                    return ptr;
            }
        }
        uint32_t emp;
        if (hint == kNoSlab && line_empty(t, L.tag, emp))
            hint = slab;
        if (L.next >= kPendingSlab) {
            if (hint == kNoSlab)
                hint = slab;
            return kNotFound;
        }
        slab = L.next;
        line_head_load(L, t.lines + (size_t)slab * kLineWords, t.tagWords);
    }
}

// Per-lane find.
__device__ inline uint32_t table_find_lane(const TableDev& t, const uint32_t* needle, uint32_t s, uint32_t h)
{
    const uint32_t tag4 = make_tag(h) * 0x01010101u;
    uint32_t slab = h % t.numBuckets;
    while (true) {
        LineHead L;
        line_head_load(L, t.lines + (size_t)slab * kLineWords, t.tagWords);
        uint32_t cand = line_match(t, L, tag4);
        while (cand) {
            const uint32_t slot = __ffs(cand) - 1u;
            cand &= cand - 1u;
            const uint32_t* item = line_item(t, slab, slot);
            bool eq = true;
            for (uint32_t k = 0; k < s && eq; ++k)
                eq = item[k] == needle[k];
            if (eq)
                return (slab << kSlotBits) | slot;
        }
        if (L.next >= kPendingSlab)
            return kNotFound;
        slab = L.next;
    }
}

// Group-cooperative insert of an item the caller knows to be absent (or, for
// bulk multiset inserts, allowed to be present).  The group's first lane
// claims an empty slot with a CAS on its tag word and the group writes the
// item; a chain that is full up to its tail grows through chain_grow.
// `start` is a line of the bucket's chain to begin at (a find's hint), or
// kNoSlab for the head.
template <int G, int MW>
__device__ inline uint32_t table_insert_group(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t h,
                                              uint32_t gl, uint32_t gbase, uint32_t start = kNoSlab)
{
    static_assert(G >= 4 && G <= 32, "group size");
    const uint32_t tag = make_tag(h);
    uint32_t slab = start == kNoSlab ? h % t.numBuckets : start;
    while (true) {
        uint32_t* line = t.lines + (size_t)slab * kLineWords;
        uint32_t ok = 0, slot = 0, next = 0, full = 0;
        if (gl == 0) {
            uint32_t tw[4];
            for (uint32_t j = 0; j < 4; ++j)
                tw[j] = j < t.tagWords ? ld_relaxed(line + j) : 0u;
            next = ld_relaxed(line + t.tagWords);
            uint32_t emp;
            line_empty(t, tw, emp);
            if (emp) {
                const uint32_t bit = __ffs(emp) - 1u;
                const uint32_t j = bit >> 2, b = bit & 3u;
                uint32_t old = 0;
                for (uint32_t q = 0; q < 4; ++q)
                    if (q == j)
                        old = tw[q];
                ok = atomicCAS(line + j, old, old | (tag << (8u * b))) == old ? 1u : 0u;
                slot = bit;
            } else {
                full = 1u;
            }
        }
        ok = group_bcast<G>(ok, 0);
        full = group_bcast<G>(full, 0);
        slot = group_bcast<G>(slot, 0);
        next = group_bcast<G>(next, 0);
        if (ok) {
            uint32_t* item = line + t.tagWords + 1u + slot * s;
#pragma unroll
            for (int k = 0; k < Needle<G, MW>::K; ++k) {
                const uint32_t j = gl + (uint32_t)k * G;
                if (j < s)
                    item[j] = n.w[k];
            }
            return (slab << kSlotBits) | slot;
        }
        if (!full || next == kPendingSlab)
            continue; // lost the CAS, or a successor is being linked: re-read the line
        if (next != kNoSlab) {
            slab = next;
            continue;
        }
        // Every group at a full tail competes for its link; the winners'
        // slabs are allocated together.
        slab = group_bcast<G>(chain_grow(t, slab, gl == 0), 0);
        if (slab == kNoSlab)
            return kNotFound;
    }
}

// Per-lane insert.
__device__ inline uint32_t table_insert_lane(const TableDev& t, const uint32_t* item, uint32_t s, uint32_t h)
{
    const uint32_t tag = make_tag(h);
    const uint32_t bucket = h % t.numBuckets;
    uint32_t slab = bucket;
    while (true) {
        uint32_t* line = t.lines + (size_t)slab * kLineWords;
        uint32_t tw[4];
        for (uint32_t j = 0; j < 4; ++j)
            tw[j] = j < t.tagWords ? ld_relaxed(line + j) : 0u;
        uint32_t emp;
        line_empty(t, tw, emp);
        if (emp) {
            const uint32_t bit = __ffs(emp) - 1u;
            const uint32_t j = bit >> 2, b = bit & 3u;
            uint32_t old = 0;
            for (uint32_t q = 0; q < 4; ++q)
                if (q == j)
                    old = tw[q];
            if (atomicCAS(line + j, old, old | (tag << (8u * b))) != old)
                continue;
            uint32_t* dst = line + t.tagWords + 1u + bit * s;
            for (uint32_t k = 0; k < s; ++k)
                dst[k] = item[k];
            return (slab << kSlotBits) | bit;
        }
        const uint32_t next = ld_relaxed(line + t.tagWords);
        if (next == kPendingSlab)
            continue; // a successor is being linked: re-read the line
        if (next != kNoSlab) {
            slab = next;
            continue;
        }
        slab = chain_grow(t, slab, true);
        if (slab == kNoSlab)
            return kNotFound;
    }
}

// ---------------------------------------------------------------- DAG encoding
struct DagDev {
    const TableDev* tables;   // [numTables], index = item words - 2
    uint32_t* fullyFilled;    // [levels][numMaterials]
    uint32_t levels;          // total levels D (voxel grid 2^D)
    uint32_t leafLevel;       // D - 2
    uint32_t M;               // material bits (0, 4 or 8)
    uint32_t numMaterials;    // 1 << M
    uint32_t minLeafWords;    // 2 + (M ? 1 : 0)
    uint32_t maxLeafWords;    // 2 + 64*M/32
    uint32_t leafSizeBits;    // bit_width(maxLeaf - minLeaf)
    uint32_t numTables;
    unsigned long long* stored; // items stored in the tables (inserts minus frees)
};

constexpr uint32_t kNodeSizeBits = 3; // bit_width(9 - 2)

__host__ __device__ inline uint32_t node_size_from_header(uint32_t header)
{
    return 1u + (uint32_t)__builtin_popcount(header & 0xFFu);
}
__host__ __device__ inline uint32_t leaf_size_from_words(uint32_t w0, uint32_t w1, uint32_t M)
{
    const uint32_t n = (uint32_t)__builtin_popcount(w0) + (uint32_t)__builtin_popcount(w1);
    return 2u + ((n * M + 31u) >> 5);
}

// Handle index <-> table pointer.  Inside the library a pointer is always
// line << 4 | slot, so every per-slot array keeps 16 entries per line.  A
// handle stores its index with only as many slot bits as the table's lines
// have slots (bit_width(line_slots(s) - 1)): 4 bits at 2-3 words, 3 at 4-5,
// 2 at 6-10, 1 at 11-15 and none from 16 words.  A table of 9-word items thus
// addresses 2^27 lines with a node handle's 29 index bits, where a fixed
// 4-bit slot code would stop it at 2^25 (2^24 at 8 material bits).  Packed
// 3 bits per size, size 2 first.
constexpr unsigned long long kSlotCodeBitsPacked =
    (4ull << 0) | (4ull << 3) | (3ull << 6) | (3ull << 9) | (2ull << 12) | (2ull << 15) | (2ull << 18)
    | (2ull << 21) | (2ull << 24) | (1ull << 27) | (1ull << 30) | (1ull << 33) | (1ull << 36) | (1ull << 39);
__host__ __device__ inline uint32_t slot_code_bits(uint32_t s)
{
    return (uint32_t)(kSlotCodeBitsPacked >> (3u * (s - 2u))) & 7u;
}
__host__ __device__ inline uint32_t ptr_to_index(uint32_t ptr, uint32_t s)
{
    return ((ptr >> kSlotBits) << slot_code_bits(s)) | (ptr & (kMaxSlots - 1u));
}
__host__ __device__ inline uint32_t index_to_ptr(uint32_t idx, uint32_t s)
{
    const uint32_t b = slot_code_bits(s);
    return ((idx >> b) << kSlotBits) | (idx & ((1u << b) - 1u));
}

__host__ __device__ inline uint32_t node_handle(uint32_t ptr, uint32_t size)
{
    return ptr_to_index(ptr, size) | ((size - 2u) << (32u - kNodeSizeBits));
}
__host__ __device__ inline uint32_t node_handle_size(uint32_t h)
{
    return 2u + (h >> (32u - kNodeSizeBits));
}
__host__ __device__ inline uint32_t node_handle_ptr(uint32_t h)
{
    return index_to_ptr(h & ((1u << (32u - kNodeSizeBits)) - 1u), node_handle_size(h));
}
__host__ __device__ inline uint32_t leaf_handle(const DagDev& d, uint32_t ptr, uint32_t size)
{
    if (d.leafSizeBits == 0)
        return ptr_to_index(ptr, size);
    return ptr_to_index(ptr, size) | ((size - d.minLeafWords) << (32u - d.leafSizeBits));
}
__host__ __device__ inline uint32_t leaf_handle_size(const DagDev& d, uint32_t h)
{
    if (d.leafSizeBits == 0)
        return d.minLeafWords;
    return d.minLeafWords + (h >> (32u - d.leafSizeBits));
}
__host__ __device__ inline uint32_t leaf_handle_ptr(const DagDev& d, uint32_t h)
{
    if (d.leafSizeBits == 0)
        return index_to_ptr(h, d.minLeafWords);
    return index_to_ptr(h & ((1u << (32u - d.leafSizeBits)) - 1u), leaf_handle_size(d, h));
}
// Most lines a table of s-word items may hold: every item's handle index
// must fit the index bits of each kind of item stored there, the index must
// not be all ones (kSentinel), and a pointer (line << 4 | slot) must fit 32
// bits.
__host__ __device__ inline uint64_t table_line_limit(uint32_t s, uint32_t minLeafWords, uint32_t maxLeafWords,
                                                     uint32_t leafSizeBits)
{
    uint32_t ib = 32u - kSlotBits + slot_code_bits(s); // pointer range, in handle-index bits
    if (s <= 9u && ib > 32u - kNodeSizeBits)
        ib = 32u - kNodeSizeBits;
    if (s >= minLeafWords && s <= maxLeafWords && ib > 32u - leafSizeBits)
        ib = 32u - leafSizeBits;
    return (1ull << (ib - slot_code_bits(s))) - 1ull;
}
__device__ inline const uint32_t* dag_node_ptr(const DagDev& d, uint32_t h)
{
    const uint32_t s = node_handle_size(h);
    return table_item_ptr(d.tables[s - 2u], node_handle_ptr(h));
}
__device__ inline const uint32_t* dag_leaf_ptr(const DagDev& d, uint32_t h)
{
    const uint32_t s = leaf_handle_size(d, h);
    return table_item_ptr(d.tables[s - 2u], leaf_handle_ptr(d, h));
}
__host__ __device__ inline bool header_full(uint32_t header, uint32_t M, uint32_t& mat)
{
    mat = M ? ((header >> 9) & ((1u << M) - 1u)) : 0u;
    return (header >> 8) & 1u;
}
__host__ __device__ inline uint32_t header_payload(uint32_t header, uint32_t M)
{
    return header >> (9u + M);
}
__host__ __device__ inline uint32_t header_make(uint32_t mask, uint32_t payload, uint32_t M)
{
    return (mask & 0xFFu) | (payload << (9u + M));
}
__host__ __device__ inline uint32_t header_make_full(uint32_t mask, uint32_t mat, uint32_t M)
{
    return (mask & 0xFFu) | (1u << 8) | (M ? (mat << 9) : 0u);
}
// get_material (my_gpu_base_dag.h:293-316), M in {0, 4, 8} (no straddling).
__host__ __device__ inline bool leaf_material(const uint32_t* leaf, uint32_t voxel, uint32_t M, uint32_t& mat)
{
    const uint64_t bits = (uint64_t)leaf[0] | ((uint64_t)leaf[1] << 32);
    if (!((bits >> voxel) & 1ull))
        return false;
    if (M == 0) {
        mat = 0;
        return true;
    }
    const uint32_t off = (uint32_t)__builtin_popcountll(bits & ((1ull << voxel) - 1ull)) * M;
    mat = (leaf[2 + (off >> 5)] >> (off & 31u)) & ((1u << M) - 1u);
    return true;
}

} // namespace svdag
