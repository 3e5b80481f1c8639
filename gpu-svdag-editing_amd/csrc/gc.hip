// Garbage collection of the editable DAG.
//
// reference: MyGPUHashDAG<GPU_Malloc>::garbageCollect (my_gpu_hash_dag.cu:364-475)
// and GpuHashTableBase::clearActiveFlags / freeInactiveItems
// (gpu_hash_table_base.cu:230-374).  Roots are the current root plus the
// level-0 fully-filled node of every material (377-380).
//
// Mark: one frontier per level, with no atomics on the items.  Every slot
// has a visit byte holding 1 + the last level the mark reached it at; a
// parent visiting a child at level L + 1 skips it if the byte already says
// L + 1, else stores L + 1 and appends it -- this replaces the reference's
// per-level radix sort + unique.  Levels run in increasing order, so the
// byte only ever grows and a byte of L + 1 can only come from the running
// level.  Two parents that visit one child at the same moment may both
// append it; the child is then expanded twice, which marks nothing new, so
// the frontier holds each (item, level) pair at least once and rarely more
// (a plain byte store also leaves the neighbouring slots' bytes alone).
// Visits must be per level: one item can occur at several levels, and the
// same child handle decodes to a node at one depth and to a leaf (another
// table) next to the leaf level.  An item is live iff its byte is nonzero.
// Sweep (always on; the reference sweeps only under CAPTURE_MEMORY_STATS_SLOW),
// in two passes: a flat pass over every line clears the tags of unmarked
// items and marks the buckets that have overflow lines; a pass over those
// chains unlinks the overflow lines left without items and pushes them on
// the slab free stack.
#include "dag_internal.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace svdag {

struct GcCounters {
    uint32_t count[kMaxLevels]; // frontier size per level (marking) / dead items per round (reference counts)
    uint32_t overflow;
    uint32_t touched;           // buckets listed for the chain pass (reference counts)
    unsigned long long itemsFreed, slabsFreed, itemsLive;
};

// Every table's descriptor, visit bytes and bucket byte map, for the
// per-table passes: one launch covers all tables, blockIdx.y picks one.
constexpr uint32_t kGcMaxTables = 20; // item sizes 2 .. 18 words use 17
struct GcTables {
    TableDev t[kGcMaxTables];
    uint8_t* visit[kGcMaxTables];
    uint8_t* hasChain[kGcMaxTables]; // per-bucket "has overflow lines" bytes, or nullptr: walk every bucket
    uint32_t* touched[kGcMaxTables]; // per-bucket bits: listed for the chain pass (zero between collections)
};

// Clears the visit bytes of the lines handed out so far (16 bytes a line).
__global__ void __launch_bounds__(256) gc_clear_visits(GcTables a)
{
    const TableDev& t = a.t[blockIdx.y];
    uint8_t* visit = a.visit[blockIdx.y];
    const uint32_t fresh = min(t.alloc->fresh, t.slabCap);
    // A pop past the free stack's end leaves its count below zero; the
    // sweep's pushes start from an empty stack.
    if (blockIdx.x == 0 && threadIdx.x == 0 && t.alloc->freeCount < 0)
        t.alloc->freeCount = 0;
    uint4* v = reinterpret_cast<uint4*>(visit);
    uint4* r = reinterpret_cast<uint4*>(t.roles);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < fresh; i += (uint64_t)gridDim.x * blockDim.x) {
        v[i] = make_uint4(0u, 0u, 0u, 0u);
        r[i] = make_uint4(0u, 0u, 0u, 0u); // the mark sets the roles items are reached in
    }
}

// Visit (item, level): true if the caller is to expand the item at L (the
// first visitor, or one of several simultaneous first visitors).  A stale
// cached byte can only read as an earlier level, which also means "expand".
__device__ inline uint8_t* gc_visit_byte(const DagDev& dag, uint8_t* const* visit, uint32_t h, bool leaf)
{
    const uint32_t s = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
    const uint32_t ptr = leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h); // slab * 16 + slot
    return &visit[s - 2][ptr];
}
__device__ inline bool gc_visit(const DagDev& dag, uint8_t* const* visit, uint32_t h, uint32_t L, bool leaf)
{
    uint8_t* b = gc_visit_byte(dag, visit, h, leaf);
    const uint8_t lv = (uint8_t)(L + 1u);
    if (*b == lv)
        return false;
    *b = lv;
    return true;
}

__global__ void gc_seed(DagDev dag, uint8_t* const* visit, const uint32_t* root, uint32_t* front, GcCounters* ctr)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    uint32_t n = 0;
    const bool leaf = dag.leafLevel == 0;
    if (*root != kSentinel && gc_visit(dag, visit, *root, 0, leaf))
        front[n++] = *root;
    for (uint32_t m = 0; m < dag.numMaterials; ++m)
        if (gc_visit(dag, visit, dag.fullyFilled[m], 0, leaf))
            front[n++] = dag.fullyFilled[m];
    ctr->count[0] = n;
}

// Expand the frontier nodes of level L: visit each child at L + 1 and append
// the interior children to be expanded to the next frontier (leaves have
// nothing to expand).  Visits use plain loads and stores, so a node's
// child visits are issued together.  Appends are staged in LDS and flushed with one atomic per
// kGcStage entries, so the frontier counter sees a few thousand atomics per
// collection instead of one per wave.
constexpr uint32_t kGcStage = 4096;

__device__ inline void gc_flush_stage(uint32_t* sStage, uint32_t& sFill, uint32_t& sBase, uint32_t* out,
                                      uint32_t outCap, uint32_t* counter, uint32_t* overflow)
{
    __syncthreads();
    const uint32_t fill = sFill;
    if (fill == 0u)
        return; // uniform: every thread read the same sFill after the barrier
    if (threadIdx.x == 0)
        sBase = atomicAdd(counter, fill);
    __syncthreads();
    const uint32_t base = sBase;
    if ((uint64_t)base + fill > outCap) {
        if (threadIdx.x == 0)
            atomicOr(overflow, 1u);
    } else {
        for (uint32_t j = threadIdx.x; j < fill; j += blockDim.x)
            out[base + j] = sStage[j];
    }
    __syncthreads();
    if (threadIdx.x == 0)
        sFill = 0u;
    __syncthreads();
}

__global__ void __launch_bounds__(256) gc_mark_level(DagDev dag, uint8_t* const* visit, uint32_t L,
                                                     const uint32_t* in, uint32_t* out, uint32_t outCap,
                                                     GcCounters* ctr)
{
    __shared__ uint32_t sStage[kGcStage];
    __shared__ uint32_t sFill, sBase, sStop;
    // After an overflow the last level's frontier has a hole: the batch
    // that crossed the end was not written, so nothing more is expanded
    // (the collection is retried with larger frontiers).  One read per
    // block: other blocks of this level may set the flag meanwhile.
    if (threadIdx.x == 0) {
        sFill = 0u;
        sStop = ctr->overflow;
    }
    __syncthreads();
    if (sStop)
        return;
    const uint32_t n = min(ctr->count[L], outCap); // both frontier buffers hold outCap entries
    const bool childLeaf = L + 1 == dag.leafLevel;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        // One pass adds at most 8 entries per thread.
        if (sFill > kGcStage - 8u * blockDim.x)
            gc_flush_stage(sStage, sFill, sBase, out, outCap, &ctr->count[L + 1], &ctr->overflow);
        const uint32_t i = i0 + threadIdx.x;
        uint32_t ch[8], keep = 0;
        if (i < n) {
            const uint32_t hn = in[i];
            const uint32_t* node = dag_node_ptr(dag, hn);
            uint8_t* role = &dag.tables[node_handle_size(hn) - 2].roles[node_handle_ptr(hn)];
            *role = *role | node_role(dag, L); // one level per kernel: equal values race
            const uint32_t cnt = __builtin_popcount(node[0] & 0xFFu);
            uint8_t* b[8];
            uint8_t v[8];
            // The node's children, then all their visit bytes, then the
            // stores: three waits per node instead of two per child.
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                ch[j] = j < cnt ? node[1 + j] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                b[j] = j < cnt ? gc_visit_byte(dag, visit, ch[j], childLeaf) : nullptr;
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                v[j] = j < cnt ? *b[j] : 0u;
            const uint8_t lv = (uint8_t)(L + 2u); // 1 + the children's level
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) {
                if (j < cnt && v[j] != lv) {
                    *b[j] = lv;
                    keep |= childLeaf ? 0u : 1u << j;
                }
            }
        }
        if (keep) {
            uint32_t pos = atomicAdd(&sFill, (uint32_t)__builtin_popcount(keep));
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if ((keep >> j) & 1u)
                    sStage[pos++] = ch[j];
        }
        __syncthreads();
    }
    gc_flush_stage(sStage, sFill, sBase, out, outCap, &ctr->count[L + 1], &ctr->overflow);
}

// Sums of per-thread counters over the block, one atomic per block (the
// counters are single words, so per-wave atomics would serialise).
__device__ inline void gc_block_add(uint32_t v, unsigned long long* dst)
{
    __shared__ uint32_t sPart[4];
    for (int d = 32; d >= 1; d >>= 1)
        v += (uint32_t)__shfl_xor((int)v, d, 64);
    __syncthreads();
    if ((threadIdx.x & 63u) == 0)
        sPart[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w)
            sum += sPart[w];
        if (sum)
            atomicAdd(dst, (unsigned long long)sum);
    }
}

// Sweep, pass 1: every line handed out so far, one lane per line (a flat,
// streaming pass).  Tags of items whose visit byte stayed zero are cleared,
// and hasChain[b] records whether bucket b has overflow lines (if the byte
// map is given).  Lines on the free stack have no tags, so they pass
// through unchanged.  blockDim.x = 256; nothing runs after a frontier
// overflow (the collection is retried with larger frontiers).
__global__ void __launch_bounds__(256) gc_sweep_lines(GcTables a, GcCounters* ctr)
{
    if (ctr->overflow)
        return;
    const TableDev& t = a.t[blockIdx.y];
    const uint8_t* visit = a.visit[blockIdx.y];
    uint8_t* hasChain = a.hasChain[blockIdx.y];
    const uint32_t fresh = min(t.alloc->fresh, t.slabCap);
    uint32_t freedItems = 0, live = 0;
    for (uint32_t slab = blockIdx.x * blockDim.x + threadIdx.x; slab < fresh; slab += gridDim.x * blockDim.x) {
        uint32_t* line = t.lines + (size_t)slab * kLineWords;
        const uint4 v = *reinterpret_cast<const uint4*>(line);
        const uint32_t tw[4] = { v.x, v.y, v.z, v.w };
        const uint4 vb = reinterpret_cast<const uint4*>(visit)[slab];
        const uint32_t vw[4] = { vb.x, vb.y, vb.z, vb.w };
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if (j >= t.tagWords)
                break;
            uint32_t keep = 0;
#pragma unroll
            for (uint32_t b = 0; b < 4; ++b) {
                const uint32_t slot = 4u * j + b;
                const uint32_t tag = (tw[j] >> (8u * b)) & 0xFFu;
                if (!tag || slot >= t.slots)
                    continue;
                if (((vw[j] >> (8u * b)) & 0xFFu) != 0u) { // slot 4j + b
                    keep |= tag << (8u * b);
                    ++live;
                } else {
                    ++freedItems;
                }
            }
            if (keep != tw[j])
                line[j] = keep;
        }
        if (hasChain && slab < t.numBuckets)
            hasChain[slab] = line[t.tagWords] != kNoSlab ? 1u : 0u;
    }
    gc_block_add(freedItems, &ctr->itemsFreed);
    gc_block_add(live, &ctr->itemsLive);
}

// Sweep, pass 2: one lane per bucket that has overflow lines (every bucket
// without the byte map) walks its chain and unlinks the overflow lines pass
// 1 left without tags; they become empty lines (link kNoSlab) on the free
// stack.  The loop is wave-uniform so the frees aggregate per wave.
__global__ void __launch_bounds__(256) gc_sweep_chains(GcTables a, GcCounters* ctr)
{
    if (ctr->overflow)
        return;
    const TableDev& t = a.t[blockIdx.y];
    const uint8_t* hasChain = a.hasChain[blockIdx.y];
    uint32_t freedSlabs = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < t.numBuckets; base += gridDim.x * blockDim.x) {
        const uint32_t b = base + threadIdx.x;
        uint32_t prev = kNoSlab, cur = kNoSlab;
        if (b < t.numBuckets && (!hasChain || hasChain[b])) {
            prev = b;
            cur = t.lines[(size_t)b * kLineWords + t.tagWords];
        }
        while (__ballot(cur != kNoSlab) != 0ull) {
            bool give = false;
            const uint32_t id = cur;
            if (cur != kNoSlab) {
                uint32_t* line = t.lines + (size_t)cur * kLineWords;
                const uint4 v = *reinterpret_cast<const uint4*>(line);
                const uint32_t next = line[t.tagWords];
                const uint32_t tags = v.x | (t.tagWords > 1 ? v.y : 0u) | (t.tagWords > 2 ? v.z : 0u)
                                      | (t.tagWords > 3 ? v.w : 0u);
                if (tags == 0u) {
                    t.lines[(size_t)prev * kLineWords + t.tagWords] = next;
                    line[t.tagWords] = kNoSlab;
                    give = true;
                    ++freedSlabs;
                } else {
                    prev = cur;
                }
                cur = next;
            }
            wave_slab_free(t, give, id);
        }
    }
    gc_block_add(freedSlabs, &ctr->slabsFreed);
}

// ---------------------------------------------------------------- recount
// After a marking collection: zero every count, then each stored item counts
// its children under each role the mark reached it in, and the level-0
// fully-filled nodes get back their permanent reference.  Nothing runs after
// a frontier overflow.
__global__ void __launch_bounds__(256) gc_refs_clear(GcTables a, GcCounters* ctr)
{
    if (ctr->overflow)
        return;
    const TableDev& t = a.t[blockIdx.y];
    const uint32_t fresh = min(t.alloc->fresh, t.slabCap);
    uint4* r = reinterpret_cast<uint4*>(t.refs); // 64 bytes a line
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)fresh * 4u;
         i += (uint64_t)gridDim.x * blockDim.x)
        r[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void __launch_bounds__(256) gc_refs_count(DagDev dag, GcTables a, GcCounters* ctr)
{
    if (ctr->overflow)
        return;
    const TableDev& t = a.t[blockIdx.y];
    const uint32_t fresh = min(t.alloc->fresh, t.slabCap);
    const uint32_t s = t.itemWords;
    for (uint32_t slab = blockIdx.x * blockDim.x + threadIdx.x; slab < fresh; slab += gridDim.x * blockDim.x) {
        const uint32_t* line = t.lines + (size_t)slab * kLineWords;
        const uint4 tv = *reinterpret_cast<const uint4*>(line);
        const uint4 rv = reinterpret_cast<const uint4*>(t.roles)[slab];
        const uint32_t tw[4] = { tv.x, tv.y, tv.z, tv.w }, rw[4] = { rv.x, rv.y, rv.z, rv.w };
        for (uint32_t slot = 0; slot < t.slots; ++slot) {
            const uint32_t sh = 8u * (slot & 3u);
            const uint32_t tag = (tw[slot >> 2] >> sh) & 0xFFu, roles = (rw[slot >> 2] >> sh) & 0xFFu;
            if (!tag || !roles)
                continue;
            const uint32_t* item = line + t.tagWords + 1u + slot * s;
            for (uint32_t c = 1; c < s; ++c) {
                if (roles & kRoleParentOfNodes)
                    child_ref_add(dag, item[c], kRoleParentOfNodes);
                if (roles & kRoleParentOfLeaves)
                    child_ref_add(dag, item[c], kRoleParentOfLeaves);
            }
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < dag.numMaterials) {
        const uint32_t h = dag.fullyFilled[threadIdx.x];
        const bool leaf = dag.leafLevel == 0;
        const uint32_t hs = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
        atomicAdd(&dag.tables[hs - 2].refs[leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h)], 1u);
    }
}

__global__ void gc_set_stored(DagDev dag, GcCounters* ctr)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && !ctr->overflow)
        *dag.stored = ctr->itemsLive;
}

// ---------------------------------------------------------------- reference-count collection
// Releases the roots held since the last collection.  A released root that
// no stored item refers to dies; every dying item releases its children
// under each of its roles, and those that drop to zero die in the next
// round (a path from a root is at most a level per round).  Dead items lose
// their tags; the overflow lines of their buckets are checked afterwards and
// unlinked when empty.  The work is proportional to what died.

// Wave-aggregated append: every lane of the wave calls it; a lane with
// n = 1 gets its index, from one atomic per wave.
__device__ inline uint32_t wave_append(uint32_t* counter, bool n)
{
    const unsigned long long m = __ballot(n);
    uint32_t base = 0;
    if (m != 0ull) {
        const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
        if (__lane_id() == leader)
            base = atomicAdd(counter, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, (int)leader, 64);
    }
    return base + (uint32_t)__popcll(m & lanes_below());
}

// Up to this many released roots travel in the first round's arguments,
// so a collection after a batch of edits needs no upload.
constexpr uint32_t kSeedInline = 256;
struct SeedInline {
    uint32_t h[kSeedInline];
};

// The table and slot of root handle h (~0u for none), as the host decodes
// the current root.
__device__ inline uint2 root_slot(const DagDev& dag, uint32_t h)
{
    if (h == kSentinel)
        return make_uint2(~0u, ~0u);
    const bool leaf = dag.leafLevel == 0;
    return make_uint2((leaf ? leaf_handle_size(dag, h) : node_handle_size(h)) - 2u,
                      leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h));
}

// What a round reads of an item before freeing it: its role, its child
// words, and for an item in an overflow line its bucket (the bucket hash of
// its words, as at insert).
struct ReleaseItem {
    uint32_t roles;
    uint32_t cw[8];
    uint32_t bucket;
    bool chained;
};

__device__ inline ReleaseItem rc_load(const DagDev& dag, bool load, uint2 e)
{
    ReleaseItem r {};
    const TableDev& t = dag.tables[e.x];
    const uint32_t s = e.x + 2u;
    const uint32_t* item = load ? table_item_ptr(t, e.y) : nullptr;
    r.roles = load ? t.roles[e.y] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
        r.cw[j] = load && j + 1u < s ? item[j + 1u] : 0u;
    r.chained = load && (e.y >> kSlotBits) >= t.numBuckets;
    if (r.chained) {
        uint32_t h = 0;
        for (uint32_t w = 0; w < s; ++w)
            h ^= hash_part(item[w], w);
        r.bucket = h % t.numBuckets;
    }
    return r;
}

// Releases item e (have) whose words v were read: frees its slot, lists an
// overflow line's bucket once for the chain pass, and takes one count from
// each child of each counted role; the children that reach zero go to the
// output list.  The slot's free and the listing are issued alongside the
// children's count loads, so each wave waits once per step: the counts, the
// decrements, the appends.  Every lane of the wave calls it.
__device__ inline void rc_release_loaded(const DagDev& dag, const GcTables& a, uint32_t rootTable, uint32_t rootPtr,
                                         bool have, uint2 e, ReleaseItem v, uint2* out, uint32_t* outCount,
                                         uint32_t cap, uint2* touchedList, GcCounters* ctr)
{
    const TableDev& t = dag.tables[e.x];
    const uint32_t s = e.x + 2u;
    const bool chained = have && v.chained;
    uint32_t tbit = 0, told = 0;
    if (chained) {
        tbit = 1u << (v.bucket & 31u);
        told = atomicOr(&a.touched[e.x][v.bucket >> 5], tbit);
    }
    if (have) {
        const uint32_t slab = e.y >> kSlotBits, slot = e.y & ((1u << kSlotBits) - 1u);
        atomicAnd(&t.lines[(size_t)slab * kLineWords + (slot >> 2)], ~(0xFFu << (8u * (slot & 3u))));
        t.roles[e.y] = 0u;
    }
    bool listed = false;
    // Children, all of one role at once (a node has at most 8): their count
    // loads, then the decrements, then one append per wave.  A role no lane
    // of the wave holds is skipped.
#pragma unroll
    for (uint32_t rb = 0; rb < 2; ++rb) {
        const uint8_t role = rb == 0 ? kRoleParentOfNodes : kRoleParentOfLeaves;
        const bool lf = role == kRoleParentOfLeaves;
        const bool mine = have && (v.roles & role);
        if (__ballot(mine) == 0ull)
            continue;
        uint2 ce[8];
        uint32_t* cp[8];
        uint32_t cv[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            cv[j] = kStickyRefs;
            if (mine && j + 1u < s) {
                const uint32_t ch = v.cw[j];
                const uint32_t cs = lf ? leaf_handle_size(dag, ch) : node_handle_size(ch);
                ce[j] = make_uint2(cs - 2u, lf ? leaf_handle_ptr(dag, ch) : node_handle_ptr(ch));
                cp[j] = &dag.tables[ce[j].x].refs[ce[j].y];
                cv[j] = ld_filter(cp[j]);
            }
        }
        if (!listed) {
            // The bucket listing rides on this round trip.
            listed = true;
            const bool list = chained && (told & tbit) == 0u;
            const uint32_t at = wave_append(&ctr->touched, list);
            if (list) {
                if (at < cap)
                    touchedList[at] = make_uint2(e.x, v.bucket);
                else
                    atomicOr(&ctr->overflow, 1u);
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j)
            if (cv[j] < kStickyRefs)
                cv[j] = atomicSub(cp[j], 1u);
        uint32_t die = 0;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            if (mine && j + 1u < s) {
                if (cv[j] == 0u)
                    atomicOr(&ctr->overflow, 1u); // counts off: the host falls back to marking
                if (cv[j] == 1u && !(ce[j].x == rootTable && ce[j].y == rootPtr))
                    die |= 1u << j;
            }
        }
        // Lanes' dying children go to consecutive slots of the round's
        // list: a wave scan of the counts, one atomic per wave.
        const uint32_t cnt = (uint32_t)__builtin_popcount(die);
        uint32_t incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t x = (uint32_t)__shfl_up((int)incl, (unsigned)d, 64);
            if (__lane_id() >= (uint32_t)d)
                incl += x;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
        uint32_t base0 = 0;
        if (total != 0u) {
            if (__lane_id() == 63u)
                base0 = atomicAdd(outCount, total);
            base0 = (uint32_t)__shfl((int)base0, 63, 64);
        }
        uint32_t at = base0 + incl - cnt;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            if ((die >> j) & 1u) {
                if (at < cap)
                    out[at] = ce[j];
                else
                    atomicOr(&ctr->overflow, 1u);
                ++at;
            }
        }
    }
    if (!listed) {
        // A wave whose items have no counted children (leaves).
        const bool list = chained && (told & tbit) == 0u;
        const uint32_t at = wave_append(&ctr->touched, list);
        if (list) {
            if (at < cap)
                touchedList[at] = make_uint2(e.x, v.bucket);
            else
                atomicOr(&ctr->overflow, 1u);
        }
    }
}

// The first round: the released roots (cand, or inl.h when cand is null)
// whose count is zero die; each root's count and words are read together.
// count[0] gets the number that died.  The counters start at zero: the
// previous collection's chain pass cleared them.
__global__ void __launch_bounds__(256) rc_release_roots(DagDev dag, GcTables a, uint32_t rootTable, uint32_t rootPtr,
                                                        const uint32_t* cand, SeedInline inl, uint32_t n, uint2* out,
                                                        uint32_t* outCount, uint32_t cap, uint2* touchedList,
                                                        GcCounters* ctr)
{
    // A pop past the free stack's end leaves its count below zero; the chain
    // pass pushes onto an empty stack then.
    if (blockIdx.x == 0 && threadIdx.x < dag.numTables && dag.tables[threadIdx.x].alloc->freeCount < 0)
        dag.tables[threadIdx.x].alloc->freeCount = 0;
    const bool leaf = dag.leafLevel == 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool in = i < n;
        uint2 e = make_uint2(0u, 0u);
        if (in) {
            const uint32_t h = cand ? cand[i] : inl.h[i];
            const uint32_t hs = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
            e = make_uint2(hs - 2u, leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h));
        }
        const uint32_t refs = in ? dag.tables[e.x].refs[e.y] : 1u;
        const ReleaseItem v = rc_load(dag, in, e);
        const bool dies = in && refs == 0u;
        const unsigned long long m = __ballot(dies);
        if (m != 0ull && __lane_id() == (uint32_t)__ffsll((long long)m) - 1u)
            atomicAdd(&ctr->count[0], (uint32_t)__popcll(m));
        rc_release_loaded(dag, a, rootTable, rootPtr, dies, e, v, out, outCount, cap, touchedList, ctr);
    }
}

// The first round of a collection inside an edit stream, enqueued behind
// the batch of edits [first, first + count) before the host has seen its
// results.  The released roots are the batch's input roots
// (status[k].rootBefore) and, for the stream's first collection, the roots
// the host held before it (inl.h[0 .. n0)), less the current root
// (*rootDev), each once.  Once an edit of the stream failed (*poison) the
// collection releases nothing: the host runs the stream again from that
// edit.  One block; n0 + count <= 256.
__global__ void __launch_bounds__(256) rc_release_batch(DagDev dag, GcTables a, const EditStatus* status,
                                                        uint32_t first, uint32_t count, SeedInline inl, uint32_t n0,
                                                        const uint32_t* rootDev, const uint32_t* poison, uint2* out,
                                                        uint32_t* outCount, uint32_t cap, uint2* touchedList,
                                                        GcCounters* ctr)
{
    __shared__ uint32_t sCand[256];
    const uint32_t t = threadIdx.x;
    const uint32_t stop = *poison;
    const uint32_t cur = *rootDev;
    const uint32_t m = n0 + count;
    uint32_t h = kSentinel;
    if (t < n0)
        h = inl.h[t];
    else if (t < m)
        h = status[first + (t - n0)].rootBefore;
    sCand[t] = h;
    // A pop past the free stack's end leaves its count below zero; the chain
    // pass pushes onto an empty stack then.
    if (!stop && t < dag.numTables && dag.tables[t].alloc->freeCount < 0)
        dag.tables[t].alloc->freeCount = 0;
    __syncthreads();
    if (stop)
        return;
    bool in = t < m && h != kSentinel && h != cur;
    for (uint32_t j = 0; j < t && in; ++j)
        in = sCand[j] != h;
    uint2 e = make_uint2(0u, 0u);
    if (in) {
        const bool leaf = dag.leafLevel == 0;
        const uint32_t hs = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
        e = make_uint2(hs - 2u, leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h));
    }
    const uint32_t refs = in ? dag.tables[e.x].refs[e.y] : 1u;
    const ReleaseItem v = rc_load(dag, in, e);
    const bool dies = in && refs == 0u;
    const unsigned long long mk = __ballot(dies);
    if (mk != 0ull && __lane_id() == (uint32_t)__ffsll((long long)mk) - 1u)
        atomicAdd(&ctr->count[0], (uint32_t)__popcll(mk));
    const uint2 rs = root_slot(dag, cur);
    rc_release_loaded(dag, a, rs.x, rs.y, dies, e, v, out, outCount, cap, touchedList, ctr);
}

// A later round: frees the items of the input list.  Each thread reads its
// first entry together with the list's length (an entry past the length is
// read but not used), so the round's chain is: entry, item, the children's
// counts, the decrements, the appends.
// rootDev (an edit stream's collections): the current root is read there
// instead of taken from rootTable / rootPtr.
__global__ void __launch_bounds__(256) rc_release(DagDev dag, GcTables a, uint32_t rootTable, uint32_t rootPtr,
                                                  const uint32_t* rootDev, const uint2* in, const uint32_t* inCount,
                                                  uint2* out, uint32_t* outCount, uint32_t cap, uint2* touchedList,
                                                  GcCounters* ctr)
{
    const uint32_t first = blockIdx.x * blockDim.x;
    const uint32_t i0 = first + threadIdx.x;
    // Once a list overflowed or a count was found off, the marking
    // collection takes over: release nothing more.
    const uint32_t stop = ctr->overflow;
    const uint32_t cnt = *inCount;
    const uint2 spec = i0 < cap ? in[i0] : make_uint2(0u, 0u);
    if (rootDev) {
        const uint2 rs = root_slot(dag, *rootDev);
        rootTable = rs.x;
        rootPtr = rs.y;
    }
    if (stop)
        return;
    const uint32_t n = min(cnt, cap);
    for (uint32_t base = first; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool have = i < n;
        const uint2 e = have ? (base == first ? spec : in[i]) : make_uint2(0u, 0u);
        const ReleaseItem v = rc_load(dag, have, e);
        rc_release_loaded(dag, a, rootTable, rootPtr, have, e, v, out, outCount, cap, touchedList, ctr);
    }
}

// The listed buckets of table blockIdx.y: unlink the overflow lines left
// without items and push them on the free stack (as gc_sweep_chains).
__global__ void __launch_bounds__(256) rc_chains(DagDev dag, GcTables a, const uint2* list, uint32_t cap,
                                                 uint32_t rounds, GcCounters* ctr, GcCounters* next, uint32_t* fail)
{
    // The release rounds have all run and each freed its whole input list:
    // the stored-item count loses the lists' sizes (once, by the first
    // block).  After an overflow the marking collection recounts.
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        unsigned long long freed = 0;
        for (uint32_t r = 0; r < rounds; ++r)
            freed += min(ctr->count[r], cap);
        ctr->itemsFreed = freed;
        *dag.stored -= freed;
        ctr->itemsLive = *dag.stored;
        // An edit stream's collection that stopped short: the host's next
        // collection marks.
        if (fail && (ctr->overflow || ctr->count[rounds] != 0u))
            *fail = 1u;
    }
    // The next counting collection's counters start at zero.
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        uint32_t* words = reinterpret_cast<uint32_t*>(next);
        for (uint32_t k = threadIdx.x; k < (uint32_t)(sizeof(GcCounters) / 4); k += blockDim.x)
            words[k] = 0u;
    }
    const uint32_t ti = blockIdx.y;
    const TableDev& t = a.t[ti];
    const uint32_t n = min(ctr->touched, cap);
    uint32_t freedSlabs = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t prev = kNoSlab, cur = kNoSlab;
        if (i < n) {
            const uint2 e = list[i];
            if (e.x == ti) {
                atomicAnd(&a.touched[ti][e.y >> 5], ~(1u << (e.y & 31u)));
                prev = e.y;
                cur = t.lines[(size_t)e.y * kLineWords + t.tagWords];
            }
        }
        while (__ballot(cur != kNoSlab) != 0ull) {
            bool give = false;
            const uint32_t id = cur;
            if (cur != kNoSlab) {
                uint32_t* line = t.lines + (size_t)cur * kLineWords;
                const uint4 v = *reinterpret_cast<const uint4*>(line);
                const uint32_t next = line[t.tagWords];
                const uint32_t tags = v.x | (t.tagWords > 1 ? v.y : 0u) | (t.tagWords > 2 ? v.z : 0u)
                                      | (t.tagWords > 3 ? v.w : 0u);
                if (tags == 0u) {
                    t.lines[(size_t)prev * kLineWords + t.tagWords] = next;
                    line[t.tagWords] = kNoSlab;
                    give = true;
                    ++freedSlabs;
                } else {
                    prev = cur;
                }
                cur = next;
            }
            wave_slab_free(t, give, id);
        }
    }
    gc_block_add(freedSlabs, &ctr->slabsFreed);
}

// Visit bytes for every table (16 per line), kept between collections;
// regrown when an arena has grown.
static int gc_masks(svdag_ctx* ctx)
{
    const size_t nt = ctx->ts.tables.size();
    bool changed = ctx->gcMasks.size() != nt;
    ctx->gcMasks.resize(nt, nullptr);
    ctx->gcMaskCap.resize(nt, 0u);
    for (size_t i = 0; i < nt; ++i) {
        const uint32_t cap = ctx->ts.tables[i].slabCap;
        if (ctx->gcMaskCap[i] >= cap && ctx->gcMasks[i])
            continue;
        if (ctx->gcMasks[i])
            (void)hipFree(ctx->gcMasks[i]);
        ctx->gcMasks[i] = nullptr;
        ctx->gcMaskCap[i] = 0;
        SVDAG_HIP(hipMalloc(&ctx->gcMasks[i], (size_t)cap * kMaxSlots));
        ctx->gcMaskCap[i] = cap;
        changed = true;
    }
    if (!ctx->dGcMasks)
        SVDAG_HIP(hipMalloc(&ctx->dGcMasks, std::max<size_t>(nt, 1) * 32 * sizeof(uint8_t*)));
    if (changed)
        SVDAG_HIP(hipMemcpyAsync(ctx->dGcMasks, ctx->gcMasks.data(), nt * sizeof(uint8_t*), hipMemcpyHostToDevice,
                                 ctx->stream));
    return SVDAG_OK;
}

// Frontier buffers (also the reference-count worklists, as uint2 pairs)
// of at least cap words, the list of buckets for the chain pass, and the
// per-table bucket bits.
static int gc_fronts(svdag_ctx* ctx, uint64_t cap)
{
    if (cap > ctx->gcCap) {
        for (int k = 0; k < 2; ++k) {
            if (ctx->gcFront[k])
                (void)hipFree(ctx->gcFront[k]);
            ctx->gcFront[k] = nullptr;
            SVDAG_HIP(hipMalloc(&ctx->gcFront[k], cap * 4));
        }
        if (ctx->dGcTouchedList)
            (void)hipFree(ctx->dGcTouchedList);
        ctx->dGcTouchedList = nullptr;
        SVDAG_HIP(hipMalloc(&ctx->dGcTouchedList, (cap / 2) * sizeof(uint2)));
        ctx->gcCap = cap;
    }
    const size_t nt = ctx->ts.tables.size();
    if (ctx->gcTouched.size() != nt) {
        for (uint32_t* p : ctx->gcTouched)
            if (p)
                (void)hipFree(p);
        ctx->gcTouched.assign(nt, nullptr);
        for (size_t i = 0; i < nt; ++i) {
            const size_t words = (ctx->ts.tables[i].numBuckets + 31u) / 32u;
            SVDAG_HIP(hipMalloc(&ctx->gcTouched[i], words * 4));
            SVDAG_HIP(hipMemsetAsync(ctx->gcTouched[i], 0, words * 4, ctx->stream));
        }
    }
    return SVDAG_OK;
}

static GcTables gc_tables(const svdag_ctx* ctx)
{
    GcTables a {};
    for (size_t i = 0; i < ctx->ts.tables.size(); ++i) {
        a.t[i] = ctx->ts.tables[i].dev;
        a.visit[i] = ctx->gcMasks[i];
        a.hasChain[i] = nullptr;
        a.touched[i] = ctx->gcTouched.size() > i ? ctx->gcTouched[i] : nullptr;
    }
    return a;
}


// The marking collection (also the fallback of the reference-count one):
// mark from the roots, sweep, then recount every stored item's references.
static int gc_mark_sweep(svdag_ctx* ctx, svdag_gc_stats* stats)
{
    const hipStream_t st = ctx->stream;
    const DagDev& dag = ctx->dag;
    // A level's frontier holds each node reached at that level once (twice
    // at most for simultaneous visits, which are rare), so the stored item
    // count sizes the first attempt; an overflow still retries larger.
    unsigned long long live = 0;
    SVDAG_HIP(hipMemcpyAsync(&live, dag.stored, sizeof(live), hipMemcpyDeviceToHost, st));
    SVDAG_HIP(hipStreamSynchronize(st));
    uint64_t cap = std::max<uint64_t>(ctx->gcCap, 1ull << 22);
    cap = std::max<uint64_t>(cap, std::min<uint64_t>(live + live / 16 + (1ull << 16), 1ull << 31));
    // SVDAG_GC_FRONTIER_ENTRIES sets the first attempt's frontier instead
    // (small values make the retries run on small DAGs, for tests).  At
    // least 4096: the seed writes the root and one node per material.
    if (const char* e = std::getenv("SVDAG_GC_FRONTIER_ENTRIES"))
        cap = std::min<uint64_t>(std::max<uint64_t>(std::strtoull(e, nullptr, 10), 4096), 1ull << 31);
    GcCounters* ctr = reinterpret_cast<GcCounters*>(ctx->gcCounters) + 2; // its own block
    const size_t nt = ctx->ts.tables.size();
    uint32_t maxCap = 1, maxBuckets = 1;
    for (size_t i = 0; i < nt; ++i) {
        maxCap = std::max(maxCap, ctx->ts.tables[i].slabCap);
        maxBuckets = std::max(maxBuckets, ctx->ts.tables[i].numBuckets);
    }
    for (;;) {
        SVDAG_TRY(gc_fronts(ctx, cap));
        // The frontiers are free after the mark; the first one holds the
        // sweep's per-bucket chain bytes, the tables' maps one after another.
        GcTables a = gc_tables(ctx);
        uint64_t mapUsed = 0;
        for (size_t i = 0; i < nt; ++i) {
            const HostTable& t = ctx->ts.tables[i];
            a.hasChain[i] = nullptr;
            if (mapUsed + t.numBuckets <= ctx->gcCap * 4) {
                a.hasChain[i] = reinterpret_cast<uint8_t*>(ctx->gcFront[0]) + mapUsed;
                mapUsed += t.numBuckets;
            }
        }
        SVDAG_HIP(hipEventRecord(ctx->ev0, st));
        SVDAG_HIP(hipMemsetAsync(ctr, 0, sizeof(GcCounters), st));
        // Bucket bits a cut-short counting collection may have left set.
        for (size_t i = 0; i < nt; ++i)
            SVDAG_HIP(hipMemsetAsync(ctx->gcTouched[i], 0, (ctx->ts.tables[i].numBuckets + 31u) / 32u * 4, st));
        gc_clear_visits<<<dim3(grid_for(maxCap, 256, 1024), (uint32_t)nt), 256, 0, st>>>(a);
        gc_seed<<<1, 64, 0, st>>>(dag, ctx->dGcMasks, ctx->dRoot, ctx->gcFront[0], ctr);
        for (uint32_t L = 0; L < dag.leafLevel; ++L)
            gc_mark_level<<<4096, 256, 0, st>>>(dag, ctx->dGcMasks, L, ctx->gcFront[L & 1],
                                                 ctx->gcFront[(L + 1) & 1], (uint32_t)cap, ctr);
        gc_sweep_lines<<<dim3(grid_for(maxCap, 256, 1024), (uint32_t)nt), 256, 0, st>>>(a, ctr);
        gc_sweep_chains<<<dim3(grid_for(maxBuckets, 256, 1024), (uint32_t)nt), 256, 0, st>>>(a, ctr);
        gc_refs_clear<<<dim3(grid_for((uint64_t)maxCap * 4, 256, 1024), (uint32_t)nt), 256, 0, st>>>(a, ctr);
        gc_refs_count<<<dim3(grid_for(maxCap, 256, 1024), (uint32_t)nt), 256, 0, st>>>(dag, a, ctr);
        gc_set_stored<<<1, 64, 0, st>>>(dag, ctr);
        SVDAG_HIP(hipEventRecord(ctx->ev1, st));
        GcCounters h {};
        SVDAG_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipStreamSynchronize(st));
        if (h.overflow) {
            // The mark ran out of frontier space and nothing after it ran:
            // run again with larger frontiers.
            if (cap >= (1ull << 31))
                break;
            cap = std::min<uint64_t>(cap * 2, 1ull << 31);
            continue;
        }
        if (stats) {
            stats->items_freed = h.itemsFreed;
            stats->slabs_freed = h.slabsFreed;
            stats->items_live = h.itemsLive;
            float ms = 0;
            SVDAG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            stats->gpu_ms = ms;
        }
        return SVDAG_OK;
    }
    set_error("garbage collection frontier did not fit");
    return SVDAG_E_CAPACITY;
}

// The reference-count collection.  done = false when it could not finish
// (a worklist overflowed, or counts were found off); the caller then runs
// the marking collection, which repairs both.
// Items per released root in each round, a decaying maximum over recent
// collections: one batch of small edits must not shrink the grids the next
// batch's large edit needs.
static void gc_take_rounds(svdag_ctx* ctx, const uint32_t* count, uint32_t rounds, size_t cand)
{
    if (cand == 0)
        return;
    ctx->gcPerRoot.resize(rounds, 0.0);
    for (uint32_t r = 0; r < rounds; ++r)
        ctx->gcPerRoot[r] = std::max((double)count[r] / (double)cand, 0.9 * ctx->gcPerRoot[r]);
}

// Takes in the counters of a collection nobody waited for.  One that
// stopped short (a list overflowed, or counts were off) left what it did
// not release stored; the next collection then marks and recounts.
int gc_settle(svdag_ctx* ctx)
{
    if (!ctx->gcPending)
        return SVDAG_OK;
    ctx->gcPending = false;
    SVDAG_HIP(hipEventSynchronize(ctx->evGc));
    const GcCounters& h = *static_cast<const GcCounters*>(ctx->hGcCtr);
    if (h.overflow || h.count[ctx->gcPendingRounds] != 0u) {
        ctx->refsDirty = true;
        return SVDAG_OK;
    }
    gc_take_rounds(ctx, h.count, ctx->gcPendingRounds, ctx->gcPendingCand);
    if (ctx->liveAtMark == 0)
        ctx->liveAtMark = h.itemsLive;
    ctx->liveAfterGc = h.itemsLive;
    return SVDAG_OK;
}

// Buffers a counting collection needs, allocated before an edit stream
// enqueues its collections.
int gc_stream_prepare(svdag_ctx* ctx)
{
    if (ctx->ts.tables.size() > kGcMaxTables) {
        set_error("too many item tables for the collection");
        return SVDAG_E_INVALID;
    }
    if (!ctx->gcCounters) {
        SVDAG_HIP(hipMalloc(&ctx->gcCounters, 3 * sizeof(GcCounters)));
        SVDAG_HIP(hipMemsetAsync(ctx->gcCounters, 0, 3 * sizeof(GcCounters), ctx->stream));
    }
    SVDAG_TRY(gc_masks(ctx));
    return gc_fronts(ctx, std::max<uint64_t>(ctx->gcCap, 1ull << 22));
}

// Enqueues the counting collection behind edits [first, first + count) of
// an edit stream (rc_release_batch): held are the roots the host held before
// the stream, for its first collection (empty for later ones), and
// held.size() + count <= 256.  *fail is set if the collection stops short.
// ctrOut: the counter block it uses.
int gc_enqueue_stream(svdag_ctx* ctx, uint32_t first, uint32_t count, const std::vector<uint32_t>& held,
                      uint32_t* fail, void** ctrOut, hipStream_t st)
{
    const DagDev& dag = ctx->dag;
    GcCounters* ctr = reinterpret_cast<GcCounters*>(ctx->gcCounters) + ctx->gcCtrIdx;
    GcCounters* next = reinterpret_cast<GcCounters*>(ctx->gcCounters) + (ctx->gcCtrIdx ^ 1u);
    ctx->gcCtrIdx ^= 1u;
    *ctrOut = ctr;
    const size_t nt = ctx->ts.tables.size();
    SeedInline inl {};
    std::copy(held.begin(), held.end(), inl.h);
    GcTables a = gc_tables(ctx);
    const uint32_t cap2 = (uint32_t)std::min<uint64_t>(ctx->gcCap / 2, 0xFFFFFFFFull);
    uint2* w[2] = { reinterpret_cast<uint2*>(ctx->gcFront[0]), reinterpret_cast<uint2*>(ctx->gcFront[1]) };
    const uint32_t rounds = dag.leafLevel + 2u;
    const double roots = (double)(count + held.size());
    auto grid = [&](uint32_t r) {
        if (r >= ctx->gcPerRoot.size())
            return 256u;
        const double want = 2.0 * ctx->gcPerRoot[r] * roots;
        return (uint32_t)std::min(256.0, std::max(1.0, std::ceil(want / 256.0)));
    };
    rc_release_batch<<<1, 256, 0, st>>>(dag, a, ctx->dStatusArr, first, count, inl, (uint32_t)held.size(), ctx->dRoot,
                                        ctx->ws.poison, w[1], &ctr->count[1], cap2, ctx->dGcTouchedList, ctr);
    for (uint32_t r = 1; r < rounds; ++r)
        rc_release<<<grid(r), 256, 0, st>>>(dag, a, ~0u, ~0u, ctx->dRoot, w[r & 1], &ctr->count[r], w[(r + 1) & 1],
                                             &ctr->count[r + 1], cap2, ctx->dGcTouchedList, ctr);
    rc_chains<<<dim3(256, (uint32_t)nt), 256, 0, st>>>(dag, a, ctx->dGcTouchedList, cap2, rounds, ctr, next, fail);
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

// The counters of an edit stream's last collection, copied to the host:
// round sizes for the grids of later collections, and the live count.
void gc_stream_take(svdag_ctx* ctx, const void* counters, uint32_t roots)
{
    const GcCounters& h = *static_cast<const GcCounters*>(counters);
    const uint32_t rounds = ctx->dag.leafLevel + 2u;
    gc_take_rounds(ctx, h.count, rounds, roots);
    if (ctx->liveAtMark == 0)
        ctx->liveAtMark = h.itemsLive;
    ctx->liveAfterGc = h.itemsLive;
}

size_t gc_counters_bytes() { return sizeof(GcCounters); }

// wait = false: the collection is only enqueued (deferred = true); its
// counters come back with the stream and gc_settle takes them in.
static int gc_release(svdag_ctx* ctx, svdag_gc_stats* stats, bool& done, bool wait, bool& deferred)
{
    done = false;
    deferred = false;
    const hipStream_t st = ctx->stream;
    const DagDev& dag = ctx->dag;
    // Two counter blocks in turn: this collection's chain pass clears the
    // other one for the next collection.
    GcCounters* ctr = reinterpret_cast<GcCounters*>(ctx->gcCounters) + ctx->gcCtrIdx;
    GcCounters* next = reinterpret_cast<GcCounters*>(ctx->gcCounters) + (ctx->gcCtrIdx ^ 1u);
    ctx->gcCtrIdx ^= 1u;
    const size_t nt = ctx->ts.tables.size();
    SVDAG_TRY(gc_fronts(ctx, std::max<uint64_t>(ctx->gcCap, 1ull << 22)));
    // Released roots: every root held since the last collection, but the
    // current one (the undo / redo history is cleared by now).
    std::vector<uint32_t> cand;
    for (uint32_t r : ctx->rootLog)
        if (r != kSentinel && r != ctx->root)
            cand.push_back(r);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    if (cand.size() > ctx->gcCandCap) {
        if (ctx->dGcCand)
            (void)hipFree(ctx->dGcCand);
        ctx->dGcCand = nullptr;
        ctx->gcCandCap = 0;
        SVDAG_HIP(hipMalloc(&ctx->dGcCand, std::max<size_t>(cand.size(), 1024) * 4));
        ctx->gcCandCap = (uint32_t)std::max<size_t>(cand.size(), 1024);
    }
    uint32_t rootTable = ~0u, rootPtr = ~0u;
    if (ctx->root != kSentinel) {
        const bool leaf = dag.leafLevel == 0;
        rootTable = (leaf ? leaf_handle_size(dag, ctx->root) : node_handle_size(ctx->root)) - 2u;
        rootPtr = leaf ? leaf_handle_ptr(dag, ctx->root) : node_handle_ptr(ctx->root);
    }
    GcTables a = gc_tables(ctx);
    const uint32_t cap2 = (uint32_t)std::min<uint64_t>(ctx->gcCap / 2, 0xFFFFFFFFull); // uint2 entries per frontier
    uint2* w[2] = { reinterpret_cast<uint2*>(ctx->gcFront[0]), reinterpret_cast<uint2*>(ctx->gcFront[1]) };
    const uint32_t rounds = dag.leafLevel + 2u;
    SVDAG_HIP(hipEventRecord(ctx->ev0, st));
    SeedInline inl {};
    const uint32_t* candDev = nullptr;
    if (cand.size() <= kSeedInline) {
        std::copy(cand.begin(), cand.end(), inl.h);
    } else {
        SVDAG_HIP(hipMemcpyAsync(ctx->dGcCand, cand.data(), cand.size() * 4, hipMemcpyHostToDevice, st));
        candDev = ctx->dGcCand;
    }
    // Each round's grid from recent collections' items per released root in
    // that round, times this collection's released roots, doubled: the
    // top rounds hold a few items per root, and a launch of 256 blocks for
    // them costs more than their work.  The grid-stride loop covers any
    // count, so the sizing only affects speed.
    auto grid = [&](uint32_t r) {
        if (r >= ctx->gcPerRoot.size() || cand.empty())
            return 256u;
        const double want = 2.0 * ctx->gcPerRoot[r] * (double)cand.size();
        return (uint32_t)std::min(256.0, std::max(1.0, std::ceil(want / 256.0)));
    };
    rc_release_roots<<<grid(0), 256, 0, st>>>(dag, a, rootTable, rootPtr, candDev, inl, (uint32_t)cand.size(), w[1],
                                              &ctr->count[1], cap2, ctx->dGcTouchedList, ctr);
    for (uint32_t r = 1; r < rounds; ++r)
        rc_release<<<grid(r), 256, 0, st>>>(dag, a, rootTable, rootPtr, nullptr, w[r & 1], &ctr->count[r],
                                             w[(r + 1) & 1], &ctr->count[r + 1], cap2, ctx->dGcTouchedList, ctr);
    rc_chains<<<dim3(256, (uint32_t)nt), 256, 0, st>>>(dag, a, ctx->dGcTouchedList, cap2, rounds, ctr, next, nullptr);
    SVDAG_HIP(hipEventRecord(ctx->ev1, st));
    if (!wait) {
        if (!ctx->hGcCtr)
            SVDAG_HIP(hipHostMalloc(&ctx->hGcCtr, sizeof(GcCounters)));
        if (!ctx->evGc)
            SVDAG_HIP(hipEventCreateWithFlags(&ctx->evGc, hipEventDisableTiming));
        SVDAG_HIP(hipMemcpyAsync(ctx->hGcCtr, ctr, sizeof(GcCounters), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipEventRecord(ctx->evGc, st));
        ctx->gcPending = true;
        ctx->gcPendingRounds = rounds;
        ctx->gcPendingCand = (uint32_t)cand.size();
        done = true;
        deferred = true;
        return SVDAG_OK;
    }
    GcCounters h {};
    SVDAG_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
    SVDAG_HIP(hipStreamSynchronize(st));
    if (h.overflow || h.count[rounds] != 0u)
        return SVDAG_OK;
    done = true;
    gc_take_rounds(ctx, h.count, rounds, cand.size());
    if (stats) {
        stats->items_freed = h.itemsFreed;
        stats->slabs_freed = h.slabsFreed;
        stats->items_live = h.itemsLive;
        float ms = 0;
        SVDAG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->gpu_ms = ms;
    }
    return SVDAG_OK;
}

int gc_run(svdag_ctx* ctx, svdag_gc_stats* stats, bool mark, bool wait)
{
    svdag_gc_stats local {};
    if (!stats)
        stats = &local;
    SVDAG_TRY(gc_settle(ctx));
    // A count that reached kStickyRefs never drops, so its item outlives
    // its last referrer until a marking collection.  One runs once the
    // stored items exceed 1.5 times the live count the last marking
    // collection found, which bounds what such items can hold.
    if (ctx->liveAtMark != 0 && ctx->liveAfterGc > ctx->liveAtMark + ctx->liveAtMark / 2)
        mark = true;
    // A collection may free the nodes of the last traced root, so the shadow
    // and occlusion passes need a new primary trace afterwards.
    ctx->traceW = ctx->traceH = 0;
    ++ctx->gcSerial;
    if (!ctx->gcCounters) {
        // Two blocks the counting collections use in turn, one for marking.
        SVDAG_HIP(hipMalloc(&ctx->gcCounters, 3 * sizeof(GcCounters)));
        SVDAG_HIP(hipMemsetAsync(ctx->gcCounters, 0, 3 * sizeof(GcCounters), ctx->stream));
    }
    if (ctx->ts.tables.size() > kGcMaxTables) {
        set_error("too many item tables for the collection");
        return SVDAG_E_INVALID;
    }
    SVDAG_TRY(gc_masks(ctx));
    bool done = false, deferred = false;
    float releaseMs = 0;
    if (!mark && !ctx->refsDirty) {
        svdag_gc_stats rs {};
        SVDAG_TRY(gc_release(ctx, &rs, done, wait, deferred));
        if (done)
            *stats = rs;
        releaseMs = rs.gpu_ms;
    }
    if (!done) {
        SVDAG_TRY(gc_mark_sweep(ctx, stats));
        stats->gpu_ms += releaseMs;
    }
    stats->marked = done ? 0u : 1u;
    if (!deferred) {
        if (!done || ctx->liveAtMark == 0)
            ctx->liveAtMark = stats->items_live;
        ctx->liveAfterGc = stats->items_live;
    }
    ctx->refsDirty = false;
    ctx->rootLog.clear();
    if (ctx->root != kSentinel)
        ctx->rootLog.push_back(ctx->root);
    return SVDAG_OK;
}

} // namespace svdag

using namespace svdag;

extern "C" int svdag_gc(svdag_ctx* ctx, svdag_gc_stats* stats)
{
    if (!ctx)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    // MyGPUHashDAGUndoRedo::garbageCollect clears both history stacks first.
    ctx->undo.clear();
    ctx->redo.clear();
    // Without stats nobody needs the collection's results now: it is only
    // enqueued on the context's stream, and what follows on that stream
    // runs after it.
    svdag_gc_stats local {};
    FrameQuiesce q(ctx);
    return gc_run(ctx, stats ? stats : &local, false, stats != nullptr);
}

extern "C" int svdag_gc_mark(svdag_ctx* ctx, svdag_gc_stats* stats)
{
    if (!ctx)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->undo.clear();
    ctx->redo.clear();
    svdag_gc_stats local {};
    FrameQuiesce q(ctx);
    return gc_run(ctx, stats ? stats : &local, true);
}
