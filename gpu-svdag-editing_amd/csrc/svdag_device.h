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
// Table layout (own design for wave64, not the reference's 32-slot slab):
//   slab  = one 128-byte tag line in a dense tag array plus a separate item
//           array [slab][120][itemWords].  The line is two 64-byte halves;
//           half h holds 15 words x 4 one-byte tags (slots 60h .. 60h+59) and
//           a copy of the next slab id in its word 15.  An item lives in the
//           half its hash selects, so a probe reads 64 bytes per slab, not
//           128.  Head slabs 0..numBuckets-1 are the buckets; overflow slabs
//           come from the slab allocator and serve both halves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svdag {

constexpr uint32_t kSlots = 120;        // item slots per slab (two halves of 60)
constexpr uint32_t kHalfSlots = 60;
constexpr uint32_t kTagWords = 32;      // words per tag line (128 B)
constexpr uint32_t kHalfWords = 16;     // words per half line (64 B)
constexpr uint32_t kHalfNext = 15;      // word of a half that holds the next slab id
constexpr uint32_t kNextWord = 31;      // the second copy (half 1); both are kept equal
constexpr uint32_t kSlotBits = 7;       // pointer = slab << 7 | slot
constexpr uint32_t kNoSlab = 0xFFFFFFFFu;
constexpr uint32_t kNotFound = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFFu;
constexpr uint32_t kMaxItemWords = 18;  // 2 + 64*8/32 (8 material bits)

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
// Half of the tag line an item lives in (independent of the bucket, h %
// numBuckets, and of the tag, the top byte of h * 0x9E3779B1).
__host__ __device__ inline uint32_t half_of(uint32_t h)
{
    return (h * 0x85EBCA6Bu) >> 31;
}
// Slot of tag byte b of word j (0..14) of half hf.
__host__ __device__ inline uint32_t slot_of(uint32_t hf, uint32_t j, uint32_t b)
{
    return hf * kHalfSlots + j * 4u + b;
}
// Tag-line word and byte of a slot.
__host__ __device__ inline uint32_t slot_word(uint32_t slot)
{
    return (slot / kHalfSlots) * kHalfWords + (slot % kHalfSlots) / 4u;
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
// Slab ids come from a lock-free free stack (64-bit head = ABA tag << 32 |
// slab id, links in freeList[id]) and, when it is empty, from the never-used
// tail of the pre-reserved arena (bump counter `fresh`).
struct AllocState {
    unsigned long long freeHead; // tag << 32 | id; id == kNoSlab when empty
    uint32_t fresh;              // next never-used slab id (starts at numBuckets)
    uint32_t oom;                // set when an allocation fails
    uint32_t freeCount;          // slabs on the free stack (for admission control)
    uint32_t pad;
};

struct TableDev {
    uint32_t* tags;      // [slabCap][32]
    uint32_t* items;     // [slabCap][kSlots][itemWords]
    uint32_t* active;    // [slabCap][4] GC mark bits
    uint32_t* freeList;  // [slabCap] free-stack links
    uint32_t* grow;      // [numBuckets] chain-growth flags
    AllocState* alloc;
    uint32_t itemWords;
    uint32_t numBuckets;
    uint32_t slabCap;
    uint32_t pad;
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

// Pop a recycled slab or bump the arena; never waits on another thread.
__device__ inline uint32_t slab_alloc(const TableDev& t)
{
    AllocState* a = t.alloc;
    unsigned long long old = ld_relaxed64(&a->freeHead);
    while ((uint32_t)old != kNoSlab) {
        const uint32_t id = (uint32_t)old;
        const uint32_t nxt = ld_relaxed(&t.freeList[id]);
        const unsigned long long nv = (((old >> 32) + 1ull) << 32) | nxt;
        const unsigned long long r = atomicCAS(&a->freeHead, old, nv);
        if (r == old) {
            atomicSub(&a->freeCount, 1u);
            return id;
        }
        old = r;
    }
    const uint32_t id = atomicAdd(&a->fresh, 1u);
    if (id >= t.slabCap) {
        atomicOr(&a->oom, 1u);
        return kNoSlab;
    }
    return id;
}

// Push a slab nobody references any more.
__device__ inline void slab_free(const TableDev& t, uint32_t id)
{
    AllocState* a = t.alloc;
    unsigned long long old = ld_relaxed64(&a->freeHead);
    while (true) {
        st_relaxed(&t.freeList[id], (uint32_t)old);
        __threadfence();
        const unsigned long long nv = (((old >> 32) + 1ull) << 32) | id;
        const unsigned long long r = atomicCAS(&a->freeHead, old, nv);
        if (r == old) {
            atomicAdd(&a->freeCount, 1u);
            return;
        }
        old = r;
    }
}

// The chain of `bucket` was seen full with head successor `observed` (read
// from the head's half-`hf` next word).  The bucket's grow flag (taken by
// CAS, never waited on) admits one grower at a time: it allocates a slab,
// clears both halves, and links it right after the head by writing both
// next copies; only flag holders move a head's successor, so the two copies
// differ only while a grower holds the flag.  Everyone else rescans.
// Returns the slab to continue with: the new slab, the head (rescan), or
// kNoSlab when the arena is exhausted.
__device__ inline uint32_t chain_grow(const TableDev& t, uint32_t bucket, uint32_t observed, uint32_t hf)
{
    uint32_t* head = t.tags + (size_t)bucket * kTagWords;
    uint32_t* mine = head + hf * kHalfWords + kHalfNext;
    if (ld_relaxed(mine) != observed)
        return bucket;
    if (atomicCAS(&t.grow[bucket], 0u, 1u) != 0u)
        return bucket;
    if (ld_relaxed(mine) != observed || ld_relaxed(head + kHalfNext) != ld_relaxed(head + kNextWord)) {
        atomicExch(&t.grow[bucket], 0u);
        return bucket;
    }
    const uint32_t id = slab_alloc(t);
    if (id == kNoSlab) {
        atomicExch(&t.grow[bucket], 0u);
        return kNoSlab;
    }
    uint32_t* nl = t.tags + (size_t)id * kTagWords;
    for (uint32_t i = 0; i < kTagWords; ++i)
        st_relaxed(nl + i, (i % kHalfWords) == kHalfNext ? observed : 0u);
    for (uint32_t i = 0; i < 4; ++i)
        t.active[(size_t)id * 4 + i] = 0u;
    __threadfence();
    atomicExch(head + kHalfNext, id);
    atomicExch(head + kNextWord, id);
    __threadfence();
    atomicExch(&t.grow[bucket], 0u);
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

// A group's view of one half line: lane gl holds words gl*HW .. gl*HW+HW-1
// of the 16 (groups wider than 16 lanes leave the upper lanes empty).
template <int G>
struct HalfLine {
    static constexpr int HW = G >= 16 ? 1 : 16 / G;
    uint32_t w[HW];
};

template <int G>
__device__ inline void halfline_load(HalfLine<G>& L, const uint32_t* half, uint32_t gl)
{
    constexpr int HW = HalfLine<G>::HW;
    if (gl * HW >= kHalfWords) {
#pragma unroll
        for (int i = 0; i < HW; ++i)
            L.w[i] = 0u;
        return;
    }
    const uint32_t* p = half + gl * HW;
    if constexpr (HW == 1) {
        L.w[0] = p[0];
    } else if constexpr (HW == 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        L.w[0] = v.x;
        L.w[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < HW / 4; ++q) {
            const uint4 v = reinterpret_cast<const uint4*>(p)[q];
            L.w[4 * q + 0] = v.x;
            L.w[4 * q + 1] = v.y;
            L.w[4 * q + 2] = v.z;
            L.w[4 * q + 3] = v.w;
        }
    }
}

// The half line's next slab id (word 15), broadcast to the group.
template <int G>
__device__ inline uint32_t halfline_next(const HalfLine<G>& L)
{
    constexpr int HW = HalfLine<G>::HW;
    return group_bcast<G>(L.w[HW - 1], kHalfNext / HW);
}

// Candidate bits (4 per word, word i of this lane at bits 4i..4i+3) of the
// slots whose tag equals tag4 (tag replicated into 4 bytes).
template <int G>
__device__ inline uint32_t halfline_match(const HalfLine<G>& L, uint32_t tag4, uint32_t gl)
{
    constexpr int HW = HalfLine<G>::HW;
    uint32_t cand = 0;
#pragma unroll
    for (int i = 0; i < HW; ++i) {
        const uint32_t j = gl * HW + i;
        if (j < kHalfNext)
            cand |= compact_bytes(zero_bytes(L.w[i] ^ tag4)) << (4 * i);
    }
    return cand;
}

// Walk a chain in half hf starting at `slab` whose half line is already in
// L; compares candidates one at a time with the whole group.
template <int G, int MW>
__device__ inline uint32_t probe_chain(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t tag4, uint32_t hf,
                                       uint32_t slab, HalfLine<G> L, uint32_t gl, uint32_t gbase)
{
    constexpr int HW = HalfLine<G>::HW;
    while (true) {
        uint32_t cand = halfline_match<G>(L, tag4, gl);
        const uint32_t next = halfline_next<G>(L);
        while (true) {
            const uint32_t bal = group_ballot<G>(cand != 0, gbase);
            if (!bal)
                break;
            const uint32_t leader = __ffs(bal) - 1u;
            uint32_t mySlot = 0;
            if (cand) {
                const uint32_t bit = __ffs(cand) - 1u;
                mySlot = slot_of(hf, gl * HW + (bit >> 2), bit & 3u);
            }
            const uint32_t slot = group_bcast<G>(mySlot, leader);
            if (gl == leader)
                cand &= cand - 1u;
            const uint32_t* item = t.items + ((size_t)slab * kSlots + slot) * s;
            const bool eq = needle_equal_lane<G>(n, item, s, gl);
            if (group_ballot<G>(!eq, gbase) == 0)
                return (slab << kSlotBits) | slot;
        }
        if (next == kNoSlab)
            return kNotFound;
        slab = next;
        halfline_load<G>(L, t.tags + (size_t)slab * kTagWords + hf * kHalfWords, gl);
    }
}

// Group-cooperative find (G in {4, 8, 16, 32}): one 64-byte half line per
// slab, read once by the group.
template <int G, int MW>
__device__ inline uint32_t table_find_group(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t h,
                                            uint32_t gl, uint32_t gbase)
{
    static_assert(G >= 4 && G <= 32, "group size");
    const uint32_t hf = half_of(h);
    const uint32_t slab = h % t.numBuckets;
    HalfLine<G> L;
    halfline_load<G>(L, t.tags + (size_t)slab * kTagWords + hf * kHalfWords, gl);
    return probe_chain<G, MW>(t, n, s, make_tag(h) * 0x01010101u, hf, slab, L, gl, gbase);
}

// Per-lane find: one lane reads the 64-byte half line with 16-byte loads.
__device__ inline uint32_t table_find_lane(const TableDev& t, const uint32_t* needle, uint32_t s, uint32_t h)
{
    const uint32_t tag4 = make_tag(h) * 0x01010101u;
    const uint32_t hf = half_of(h);
    uint32_t slab = h % t.numBuckets;
    while (true) {
        const uint4* line = reinterpret_cast<const uint4*>(t.tags + (size_t)slab * kTagWords + hf * kHalfWords);
        uint32_t next = kNoSlab;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = line[q];
            const uint32_t ws[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t j = (uint32_t)(q * 4 + i);
                if (j == kHalfNext) {
                    next = ws[i];
                    continue;
                }
                uint32_t c = compact_bytes(zero_bytes(ws[i] ^ tag4));
                while (c) {
                    const uint32_t slot = slot_of(hf, j, __ffs(c) - 1u);
                    c &= c - 1u;
                    const uint32_t* item = t.items + ((size_t)slab * kSlots + slot) * s;
                    bool eq = true;
                    for (uint32_t k = 0; k < s && eq; ++k)
                        eq = item[k] == needle[k];
                    if (eq)
                        return (slab << kSlotBits) | slot;
                }
            }
        }
        if (next == kNoSlab)
            return kNotFound;
        slab = next;
    }
}

// Group-cooperative insert of an item the caller knows to be absent (or, for
// bulk multiset inserts, allowed to be present).  Claims a slot of the
// item's half with a CAS on its tag word; a chain whose halves `hf` are all
// full grows through chain_grow.
template <int G, int MW>
__device__ inline uint32_t table_insert_group(const TableDev& t, const Needle<G, MW>& n, uint32_t s, uint32_t h,
                                              uint32_t gl, uint32_t gbase)
{
    static_assert(G >= 4 && G <= 32, "group size");
    constexpr int HW = HalfLine<G>::HW;
    const uint32_t tag = make_tag(h);
    const uint32_t hf = half_of(h);
    const uint32_t bucket = h % t.numBuckets;
    uint32_t slab = bucket;
    uint32_t headNext = kNoSlab;
    while (true) {
        uint32_t* half = t.tags + (size_t)slab * kTagWords + hf * kHalfWords;
        HalfLine<G> L;
        if (gl * HW < kHalfWords) {
#pragma unroll
            for (int i = 0; i < HW; ++i)
                L.w[i] = ld_relaxed(half + gl * HW + i);
        } else {
#pragma unroll
            for (int i = 0; i < HW; ++i)
                L.w[i] = 0u;
        }
        uint32_t emp = 0;
#pragma unroll
        for (int i = 0; i < HW; ++i) {
            const uint32_t j = gl * HW + i;
            if (j < kHalfNext)
                emp |= compact_bytes(zero_bytes(L.w[i])) << (4 * i);
        }
        const uint32_t next = halfline_next<G>(L);
        const uint32_t bal = group_ballot<G>(emp != 0, gbase);
        if (bal) {
            const uint32_t leader = __ffs(bal) - 1u;
            uint32_t ok = 0, slot = 0;
            if (gl == leader) {
                const uint32_t bit = __ffs(emp) - 1u;
                const uint32_t i = bit >> 2, b = bit & 3u;
                uint32_t old = 0;
#pragma unroll
                for (int q = 0; q < HW; ++q)
                    if ((uint32_t)q == i)
                        old = L.w[q];
                const uint32_t r = atomicCAS(half + gl * HW + i, old, old | (tag << (8u * b)));
                ok = (r == old) ? 1u : 0u;
                slot = slot_of(hf, gl * HW + i, b);
            }
            ok = group_bcast<G>(ok, leader);
            slot = group_bcast<G>(slot, leader);
            if (ok) {
                uint32_t* item = t.items + ((size_t)slab * kSlots + slot) * s;
#pragma unroll
                for (int k = 0; k < Needle<G, MW>::K; ++k) {
                    const uint32_t j = gl + (uint32_t)k * G;
                    if (j < s)
                        item[j] = n.w[k];
                }
                return (slab << kSlotBits) | slot;
            }
            continue;
        }
        if (slab == bucket)
            headNext = next;
        if (next != kNoSlab) {
            slab = next;
            continue;
        }
        // One group per wave grows a chain per round; the others rescan.
        const unsigned long long growers = __ballot(gl == 0);
        uint32_t grown = bucket;
        if (gl == 0 && (uint32_t)(__ffsll((long long)growers) - 1) == __lane_id())
            grown = chain_grow(t, bucket, headNext, hf);
        slab = group_bcast<G>(grown, 0);
        if (slab == kNoSlab)
            return kNotFound;
    }
}

// Per-lane insert.
__device__ inline uint32_t table_insert_lane(const TableDev& t, const uint32_t* item, uint32_t s, uint32_t h)
{
    const uint32_t tag = make_tag(h);
    const uint32_t hf = half_of(h);
    const uint32_t bucket = h % t.numBuckets;
    uint32_t slab = bucket;
    uint32_t headNext = kNoSlab;
    while (true) {
        uint32_t* half = t.tags + (size_t)slab * kTagWords + hf * kHalfWords;
        bool retry = false;
        for (uint32_t w = 0; w < kHalfNext; ++w) {
            const uint32_t v = ld_relaxed(half + w);
            uint32_t e = compact_bytes(zero_bytes(v));
            if (!e)
                continue;
            const uint32_t b = __ffs(e) - 1u;
            const uint32_t r = atomicCAS(half + w, v, v | (tag << (8u * b)));
            if (r != v) {
                retry = true;
                break;
            }
            const uint32_t slot = slot_of(hf, w, b);
            uint32_t* dst = t.items + ((size_t)slab * kSlots + slot) * s;
            for (uint32_t j = 0; j < s; ++j)
                dst[j] = item[j];
            return (slab << kSlotBits) | slot;
        }
        if (retry)
            continue;
        const uint32_t next = ld_relaxed(half + kHalfNext);
        if (slab == bucket)
            headNext = next;
        if (next != kNoSlab) {
            slab = next;
            continue;
        }
        // One lane per wave grows a chain per round; the others rescan.
        const unsigned long long growers = __ballot(1);
        slab = bucket;
        if ((uint32_t)(__ffsll((long long)growers) - 1) == __lane_id())
            slab = chain_grow(t, bucket, headNext, hf);
        if (slab == kNoSlab)
            return kNotFound;
    }
}

__host__ __device__ inline const uint32_t* table_item_ptr(const TableDev& t, uint32_t ptr)
{
    const uint32_t slab = ptr >> kSlotBits, slot = ptr & ((1u << kSlotBits) - 1u);
    return t.items + ((size_t)slab * kSlots + slot) * t.itemWords;
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
__host__ __device__ inline uint32_t node_handle(uint32_t ptr, uint32_t size)
{
    return ptr | ((size - 2u) << (32u - kNodeSizeBits));
}
__host__ __device__ inline uint32_t node_handle_size(uint32_t h)
{
    return 2u + (h >> (32u - kNodeSizeBits));
}
__host__ __device__ inline uint32_t node_handle_ptr(uint32_t h)
{
    return h & ((1u << (32u - kNodeSizeBits)) - 1u);
}
__host__ __device__ inline uint32_t leaf_handle(const DagDev& d, uint32_t ptr, uint32_t size)
{
    if (d.leafSizeBits == 0)
        return ptr;
    return ptr | ((size - d.minLeafWords) << (32u - d.leafSizeBits));
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
        return h;
    return h & ((1u << (32u - d.leafSizeBits)) - 1u);
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
