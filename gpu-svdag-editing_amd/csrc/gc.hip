// Garbage collection of the editable DAG.
//
// reference: MyGPUHashDAG<GPU_Malloc>::garbageCollect (my_gpu_hash_dag.cu:364-475)
// and GpuHashTableBase::clearActiveFlags / freeInactiveItems
// (gpu_hash_table_base.cu:230-374).  Roots are the current root plus the
// level-0 fully-filled node of every material (377-380).
//
// Mark: one frontier per level holding each (item, level) pair once.  Every
// slot has a 32-bit mask of the levels it was reached at; a parent claims a
// child for level L + 1 with the atomicOr that sets bit L + 1 (after a
// relaxed load shows it unset), and only the claimer appends it -- this
// replaces the reference's per-level radix sort + unique.  Deduplication must be per
// level: one item can occur at several levels, and the same child handle
// decodes to a node at one depth and to a leaf (another table) next to the
// leaf level.
// Sweep (always on; the reference sweeps only under CAPTURE_MEMORY_STATS_SLOW):
// tags of unmarked items are cleared and wholly unmarked overflow slabs are
// unlinked and pushed on the slab free stack.
#include "dag_internal.h"
#include <algorithm>

namespace svdag {

struct GcCounters {
    uint32_t count[kMaxLevels]; // frontier size per level
    uint32_t overflow;
    uint32_t pad;
    unsigned long long itemsFreed, slabsFreed, itemsLive;
};

__global__ void gc_clear_active(TableDev t, uint32_t* levelMask)
{
    const uint32_t fresh = min(t.alloc->fresh, t.slabCap);
    // A pop past the free stack's end leaves its count below zero; the
    // sweep's pushes start from an empty stack.
    if (blockIdx.x == 0 && threadIdx.x == 0 && t.alloc->freeCount < 0)
        t.alloc->freeCount = 0;
    const uint64_t n = (uint64_t)fresh * kMaxSlots;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        levelMask[i] = 0u;
        if (i < fresh)
            t.active[i] = 0u;
    }
}

// Claim (item, level): true for exactly one caller, the first to reach the
// item at that level.  A relaxed load filters the common case of an item
// already claimed (DAG children are shared by many parents), so only first
// visits pay for the atomic.  The first claim at any level also sets the
// item's active bit.
__device__ inline bool gc_claim(const DagDev& dag, uint32_t* const* levelMask, uint32_t h, uint32_t L, bool leaf)
{
    const uint32_t s = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
    const uint32_t ptr = leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h);
    const uint32_t slab = ptr >> kSlotBits, slot = ptr & ((1u << kSlotBits) - 1u);
    uint32_t* m = &levelMask[s - 2][(size_t)slab * kMaxSlots + slot];
    const uint32_t bit = 1u << L;
    if (ld_relaxed(m) & bit)
        return false;
    const uint32_t seen = atomicOr(m, bit);
    if (seen & bit)
        return false;
    if (!seen)
        atomicOr(&dag.tables[s - 2].active[slab], 1u << slot);
    return true;
}

__global__ void gc_seed(DagDev dag, uint32_t* const* levelMask, const uint32_t* root, uint32_t* front,
                        GcCounters* ctr)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    uint32_t n = 0;
    const bool leaf = dag.leafLevel == 0;
    if (*root != kSentinel && gc_claim(dag, levelMask, *root, 0, leaf))
        front[n++] = *root;
    for (uint32_t m = 0; m < dag.numMaterials; ++m)
        if (gc_claim(dag, levelMask, dag.fullyFilled[m], 0, leaf))
            front[n++] = dag.fullyFilled[m];
    ctr->count[0] = n;
}

// Expand the claimed nodes of level L: claim each child at L + 1 and append
// the claimed interior children to the next frontier (leaves have nothing to
// expand).  Appends are reserved once per wave.
__global__ void __launch_bounds__(256) gc_mark_level(DagDev dag, uint32_t* const* levelMask, uint32_t L,
                                                     const uint32_t* in, uint32_t* out, uint32_t outCap,
                                                     GcCounters* ctr)
{
    const uint32_t n = min(ctr->count[L], outCap); // both frontier buffers hold outCap entries
    const bool childLeaf = L + 1 == dag.leafLevel;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waveBase = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint32_t i0 = waveBase; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + lane;
        uint32_t kept[8];
        uint32_t k = 0;
        if (i < n) {
            const uint32_t* node = dag_node_ptr(dag, in[i]);
            const uint32_t cnt = __builtin_popcount(node[0] & 0xFFu);
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t ch = node[1 + j];
                if (gc_claim(dag, levelMask, ch, L + 1, childLeaf) && !childLeaf)
                    kept[k++] = ch;
            }
        }
        uint32_t incl = k;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)incl, d, 64);
            if (lane >= d)
                incl += v;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
        if (!total)
            continue;
        uint32_t base = 0;
        if (lane == 63)
            base = atomicAdd(&ctr->count[L + 1], total);
        base = (uint32_t)__shfl((int)base, 63, 64);
        if ((uint64_t)base + total > outCap) {
            if (lane == 0)
                atomicOr(&ctr->overflow, 1u);
            continue;
        }
        const uint32_t off = base + incl - k;
        for (uint32_t j = 0; j < k; ++j)
            out[off + j] = kept[j];
    }
}

// Thirty-two lanes per bucket chain, two chains per wave: lane w holds word w
// of the slab's line (one coalesced 128-byte read); the tag-word lanes clear
// the tags of unmarked items, and the chain link (word 31) is broadcast to
// walk on.  Overflow slabs with no marked item are unlinked and freed.  The
// loop over chains is wave-uniform so the cross-lane operations see every
// lane.
__global__ void __launch_bounds__(256) gc_sweep(TableDev t, GcCounters* ctr)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = lane & 31u;
    const uint32_t half = lane >> 5;
    const bool tagWord = w < t.tagWords;
    const uint32_t chainsPerBlock = blockDim.x / 32u;
    const uint32_t stride = gridDim.x * chainsPerBlock;
    uint32_t freedItems = 0, freedSlabs = 0, live = 0;
    for (uint32_t c0 = blockIdx.x * chainsPerBlock + ((threadIdx.x >> 6) << 1); c0 < t.numBuckets; c0 += stride) {
        const uint32_t b0 = c0 + half;
        uint32_t slab = b0 < t.numBuckets ? b0 : kNoSlab;
        uint32_t prev = kNoSlab;
        while (__ballot(slab != kNoSlab) != 0ull) {
            uint32_t v = 0, keep = 0;
            if (slab != kNoSlab) {
                uint32_t* line = t.lines + (size_t)slab * kLineWords;
                v = line[w];
                if (tagWord) {
                    const uint32_t a = t.active[slab] >> (4u * w);
                    for (uint32_t byte = 0; byte < 4; ++byte) {
                        const uint32_t tag = (v >> (8 * byte)) & 0xFFu;
                        if (!tag || 4u * w + byte >= t.slots)
                            continue;
                        if ((a >> byte) & 1u) {
                            keep |= tag << (8 * byte);
                            ++live;
                        } else {
                            ++freedItems;
                        }
                    }
                    if (keep != v)
                        line[w] = keep;
                }
            }
            const uint32_t next = (uint32_t)__shfl((int)v, (int)(half * 32u + kNextWord), 64);
            const unsigned long long kept = __ballot(tagWord && keep != 0u);
            const bool anyKept = ((kept >> (half * 32u)) & 0xFFFFFFFFull) != 0ull;
            const bool unlink = slab != kNoSlab && slab != b0 && !anyKept;
            if (unlink && w == 0) {
                t.lines[(size_t)prev * kLineWords + kNextWord] = next;
                t.lines[(size_t)slab * kLineWords + kNextWord] = kNoSlab; // tags were cleared above
                ++freedSlabs;
            }
            wave_slab_free(t, unlink && w == 0, slab);
            if (slab != kNoSlab) {
                if (!unlink)
                    prev = slab;
                slab = next;
            }
        }
    }
    // Wave sums, then one atomic per counter per wave.
    for (int m = 32; m >= 1; m >>= 1) {
        freedItems += (uint32_t)__shfl_xor((int)freedItems, m, 64);
        freedSlabs += (uint32_t)__shfl_xor((int)freedSlabs, m, 64);
        live += (uint32_t)__shfl_xor((int)live, m, 64);
    }
    if (lane == 0) {
        if (freedItems)
            atomicAdd(&ctr->itemsFreed, (unsigned long long)freedItems);
        if (freedSlabs)
            atomicAdd(&ctr->slabsFreed, (unsigned long long)freedSlabs);
        if (live)
            atomicAdd(&ctr->itemsLive, (unsigned long long)live);
    }
}

// Level masks for every table, kept between collections; regrown when an
// arena has grown.
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
        SVDAG_HIP(hipMalloc(&ctx->gcMasks[i], (size_t)cap * kMaxSlots * 4));
        ctx->gcMaskCap[i] = cap;
        changed = true;
    }
    if (!ctx->dGcMasks)
        SVDAG_HIP(hipMalloc(&ctx->dGcMasks, std::max<size_t>(nt, 1) * 32 * sizeof(uint32_t*)));
    if (changed)
        SVDAG_HIP(hipMemcpyAsync(ctx->dGcMasks, ctx->gcMasks.data(), nt * sizeof(uint32_t*), hipMemcpyHostToDevice,
                                 ctx->stream));
    return SVDAG_OK;
}

int gc_run(svdag_ctx* ctx, svdag_gc_stats* stats)
{
    const hipStream_t st = ctx->stream;
    const DagDev& dag = ctx->dag;
    uint64_t cap = std::max<uint64_t>(ctx->gcCap, 1ull << 22);
    if (!ctx->gcCounters)
        SVDAG_HIP(hipMalloc(&ctx->gcCounters, sizeof(GcCounters)));
    GcCounters* ctr = reinterpret_cast<GcCounters*>(ctx->gcCounters);
    SVDAG_TRY(gc_masks(ctx));
    const size_t nt = ctx->ts.tables.size();
    for (int attempt = 0; attempt < 6; ++attempt) {
        if (cap > ctx->gcCap) {
            for (int k = 0; k < 2; ++k) {
                if (ctx->gcFront[k])
                    (void)hipFree(ctx->gcFront[k]);
                ctx->gcFront[k] = nullptr;
                SVDAG_HIP(hipMalloc(&ctx->gcFront[k], cap * 4));
            }
            ctx->gcCap = cap;
        }
        SVDAG_HIP(hipEventRecord(ctx->ev0, st));
        SVDAG_HIP(hipMemsetAsync(ctr, 0, sizeof(GcCounters), st));
        for (size_t i = 0; i < nt; ++i)
            gc_clear_active<<<grid_for((uint64_t)ctx->ts.tables[i].slabCap * kMaxSlots, 256, 8192), 256, 0, st>>>(
                ctx->ts.tables[i].dev, ctx->gcMasks[i]);
        gc_seed<<<1, 64, 0, st>>>(dag, ctx->dGcMasks, ctx->dRoot, ctx->gcFront[0], ctr);
        for (uint32_t L = 0; L < dag.leafLevel; ++L)
            gc_mark_level<<<4096, 256, 0, st>>>(dag, ctx->dGcMasks, L, ctx->gcFront[L & 1],
                                                 ctx->gcFront[(L + 1) & 1],
                                                 (uint32_t)std::min<uint64_t>(ctx->gcCap, 0xFFFFFFFFull), ctr);
        GcCounters h {};
        SVDAG_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipStreamSynchronize(st));
        if (h.overflow) {
            cap = std::min<uint64_t>(cap * 2, 1ull << 31);
            continue;
        }
        for (auto& t : ctx->ts.tables)
            gc_sweep<<<grid_for(t.numBuckets, 8, 8192), 256, 0, st>>>(t.dev, ctr);
        SVDAG_HIP(hipEventRecord(ctx->ev1, st));
        SVDAG_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipStreamSynchronize(st));
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
    svdag_gc_stats local {};
    FrameQuiesce q(ctx);
    return gc_run(ctx, stats ? stats : &local);
}
