// Garbage collection of the editable DAG.
//
// reference: MyGPUHashDAG<GPU_Malloc>::garbageCollect (my_gpu_hash_dag.cu:364-475)
// and GpuHashTableBase::clearActiveFlags / freeInactiveItems
// (gpu_hash_table_base.cu:230-374).  Roots are the current root plus the
// level-0 fully-filled node of every material (377-380).
//
// Mark: one frontier per level.  Every slot has a 32-bit mask of the levels
// it was reached at; the atomicOr that sets bit L returns the old mask, so
// only the first visitor of an (item, level) pair expands it -- this replaces
// the reference's per-level radix sort + unique.  Deduplication must be per
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
    const uint64_t n = (uint64_t)fresh * kSlots;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        levelMask[i] = 0u;
        if (i < (uint64_t)fresh * 4)
            t.active[i] = 0u;
    }
}

__global__ void gc_seed(DagDev dag, const uint32_t* root, uint32_t* front, GcCounters* ctr)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    uint32_t n = 0;
    if (*root != kSentinel)
        front[n++] = *root;
    for (uint32_t m = 0; m < dag.numMaterials; ++m)
        front[n++] = dag.fullyFilled[m];
    ctr->count[0] = n;
}

__global__ void gc_mark_level(DagDev dag, uint32_t* const* levelMask, uint32_t L, const uint32_t* in, uint32_t* out,
                              uint32_t outCap, GcCounters* ctr)
{
    const uint32_t n = min(ctr->count[L], outCap); // both frontier buffers hold outCap entries
    const bool leaf = L == dag.leafLevel;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t h = in[i];
        const uint32_t s = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
        const uint32_t ptr = leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h);
        const TableDev& t = dag.tables[s - 2];
        const uint32_t slab = ptr >> kSlotBits, slot = ptr & ((1u << kSlotBits) - 1u);
        const uint32_t seen = atomicOr(&levelMask[s - 2][(size_t)slab * kSlots + slot], 1u << L);
        if (seen & (1u << L))
            continue;
        if (!seen)
            atomicOr(&t.active[(size_t)slab * 4 + (slot >> 5)], 1u << (slot & 31u));
        if (leaf)
            continue;
        const uint32_t* node = table_item_ptr(t, ptr);
        const uint32_t mask = node[0] & 0xFFu;
        const uint32_t k = __builtin_popcount(mask);
        const uint32_t base = atomicAdd(&ctr->count[L + 1], k);
        if ((uint64_t)base + k > outCap) {
            atomicOr(&ctr->overflow, 1u);
            continue;
        }
        for (uint32_t j = 0; j < k; ++j)
            out[base + j] = node[1 + j];
    }
}

// One lane per bucket walks the chain.
__global__ void gc_sweep(TableDev t, GcCounters* ctr)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < t.numBuckets; b += gridDim.x * blockDim.x) {
        uint32_t slab = b;
        uint32_t prev = kNoSlab;
        uint32_t freedItems = 0, freedSlabs = 0, live = 0;
        while (slab != kNoSlab) {
            uint32_t* line = t.tags + (size_t)slab * kTagWords;
            const uint32_t* act = t.active + (size_t)slab * 4;
            uint32_t anyActive = 0;
            for (uint32_t w = 0; w < kNextWord; ++w) {
                const uint32_t v = line[w];
                if (!v)
                    continue;
                const uint32_t a = (act[(w * 4) >> 5] >> ((w * 4) & 31u)) & 0xFu;
                uint32_t keep = 0;
                for (uint32_t byte = 0; byte < 4; ++byte) {
                    const uint32_t tag = (v >> (8 * byte)) & 0xFFu;
                    if (!tag)
                        continue;
                    if ((a >> byte) & 1u) {
                        keep |= tag << (8 * byte);
                        ++live;
                    } else {
                        ++freedItems;
                    }
                }
                line[w] = keep;
                anyActive |= keep;
            }
            const uint32_t next = line[kNextWord];
            if (slab != b && anyActive == 0) {
                t.tags[(size_t)prev * kTagWords + kNextWord] = next;
                slab_free(t, slab);
                ++freedSlabs;
            } else {
                prev = slab;
            }
            slab = next;
        }
        if (freedItems)
            atomicAdd(&ctr->itemsFreed, (unsigned long long)freedItems);
        if (freedSlabs)
            atomicAdd(&ctr->slabsFreed, (unsigned long long)freedSlabs);
        if (live)
            atomicAdd(&ctr->itemsLive, (unsigned long long)live);
    }
}

int gc_run(svdag_ctx* ctx, svdag_gc_stats* stats)
{
    const hipStream_t st = ctx->stream;
    const DagDev& dag = ctx->dag;
    uint64_t cap = std::max<uint64_t>(ctx->gcCap, 1ull << 22);
    if (!ctx->gcCounters)
        SVDAG_HIP(hipMalloc(&ctx->gcCounters, sizeof(GcCounters)));
    GcCounters* ctr = reinterpret_cast<GcCounters*>(ctx->gcCounters);
    // Per-slot level masks, sized to each table's current slab capacity.
    const size_t nt = ctx->ts.tables.size();
    std::vector<uint32_t*> masks(nt, nullptr);
    uint32_t** dMasks = nullptr;
    for (size_t i = 0; i < nt; ++i)
        SVDAG_HIP(hipMalloc(&masks[i], (size_t)ctx->ts.tables[i].slabCap * kSlots * 4));
    SVDAG_HIP(hipMalloc(&dMasks, nt * sizeof(uint32_t*)));
    SVDAG_HIP(hipMemcpyAsync(dMasks, masks.data(), nt * sizeof(uint32_t*), hipMemcpyHostToDevice, st));
    struct Release {
        std::vector<uint32_t*>& m;
        uint32_t**& d;
        ~Release()
        {
            for (auto* p : m)
                if (p)
                    (void)hipFree(p);
            if (d)
                (void)hipFree(d);
        }
    } release { masks, dMasks };
    for (int attempt = 0; attempt < 6; ++attempt) {
        if (cap > ctx->gcCap) {
            for (int k = 0; k < 2; ++k) {
                if (ctx->gcFront[k])
                    (void)hipFree(ctx->gcFront[k]);
                SVDAG_HIP(hipMalloc(&ctx->gcFront[k], cap * 4));
            }
            ctx->gcCap = cap;
        }
        SVDAG_HIP(hipEventRecord(ctx->ev0, st));
        SVDAG_HIP(hipMemsetAsync(ctr, 0, sizeof(GcCounters), st));
        for (size_t i = 0; i < nt; ++i)
            gc_clear_active<<<grid_for((uint64_t)ctx->ts.tables[i].slabCap * kSlots, 256, 8192), 256, 0, st>>>(
                ctx->ts.tables[i].dev, masks[i]);
        gc_seed<<<1, 64, 0, st>>>(dag, ctx->dRoot, ctx->gcFront[0], ctr);
        for (uint32_t L = 0; L <= dag.leafLevel; ++L)
            gc_mark_level<<<4096, 256, 0, st>>>(dag, dMasks, L, ctx->gcFront[L & 1], ctx->gcFront[(L + 1) & 1],
                                                 (uint32_t)std::min<uint64_t>(ctx->gcCap, 0xFFFFFFFFull), ctr);
        GcCounters h {};
        SVDAG_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipStreamSynchronize(st));
        if (h.overflow) {
            cap = std::min<uint64_t>(cap * 2, 1ull << 31);
            continue;
        }
        for (auto& t : ctx->ts.tables)
            gc_sweep<<<grid_for(t.numBuckets, 256, 8192), 256, 0, st>>>(t.dev, ctr);
        SVDAG_HIP(hipEventRecord(ctx->ev1, st));
        SVDAG_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipStreamSynchronize(st));
        (void)h;
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
