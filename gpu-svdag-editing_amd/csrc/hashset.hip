// Slab hash tables on the device: allocation, initialisation, growth and the
// bulk insert / find kernels (per-lane and wave64-group variants).
//
// Stands in for CompactAccelerationHashTable / GpuHashTableBase
// (hash_tables/compact_acceleration_hash_table.{h,cu}, gpu_hash_table_base.{h,cu})
// and the HashMapBench bulk kernels (extras/hashmap_bench/src/main.cu:124-216).
#include "svdag_internal.h"
#include <cmath>
#include <algorithm>
#include <cstring>
#include <mutex>

namespace svdag {

static thread_local std::string g_lastError;
void set_error(const std::string& msg) { g_lastError = msg; }

// Head slabs: no tags, no successor, no active bits.
// Lines [first, last) become empty: tag words 0, next kNoSlab.
__global__ void table_init_lines(TableDev t, uint32_t first, uint32_t last)
{
    const size_t n = (size_t)(last - first) * kLineWords;
    uint32_t* base = t.lines + (size_t)first * kLineWords;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t w = (uint32_t)(i % kLineWords);
        if (w == t.tagWords)
            base[i] = kNoSlab;
        else if (w < t.tagWords)
            base[i] = 0u;
    }
}

int table_alloc(HostTable& t, uint32_t itemWords, uint32_t numBuckets, uint32_t slabCap, hipStream_t stream,
                bool withRefs)
{
    if (itemWords == 0 || itemWords > kMaxItemWords || numBuckets == 0 || slabCap < numBuckets) {
        set_error("table_alloc: bad size");
        return SVDAG_E_INVALID;
    }
    if (((uint64_t)slabCap << kSlotBits) > (1ull << 32)) {
        set_error("table_alloc: slab capacity exceeds the pointer range");
        return SVDAG_E_INVALID;
    }
    t.itemWords = itemWords;
    t.numBuckets = numBuckets;
    t.slabCap = slabCap;
    TableDev& d = t.dev;
    d.itemWords = itemWords;
    d.numBuckets = numBuckets;
    d.slabCap = slabCap;
    d.slots = line_slots(itemWords);
    d.tagWords = line_tag_words(d.slots);
    SVDAG_HIP(hipMalloc(&d.lines, (size_t)slabCap * kLineWords * 4));
    SVDAG_HIP(hipMalloc(&d.freeList, (size_t)slabCap * 4));
    SVDAG_HIP(hipMalloc(&d.alloc, sizeof(AllocState)));
    if (withRefs) {
        SVDAG_HIP(hipMalloc(&d.refs, (size_t)slabCap * kMaxSlots * 4));
        SVDAG_HIP(hipMalloc(&d.roles, (size_t)slabCap * kMaxSlots));
        SVDAG_HIP(hipMemsetAsync(d.refs, 0, (size_t)slabCap * kMaxSlots * 4, stream));
        SVDAG_HIP(hipMemsetAsync(d.roles, 0, (size_t)slabCap * kMaxSlots, stream));
    }
    AllocState a { numBuckets, 0, 0u, 0u };
    SVDAG_HIP(hipMemcpyAsync(d.alloc, &a, sizeof(a), hipMemcpyHostToDevice, stream));
    table_init_lines<<<grid_for((uint64_t)slabCap * kLineWords, 256, 8192), 256, 0, stream>>>(d, 0u, slabCap);
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

void table_free(HostTable& t)
{
    TableDev& d = t.dev;
    if (d.lines) (void)hipFree(d.lines);
    if (d.freeList) (void)hipFree(d.freeList);
    if (d.alloc) (void)hipFree(d.alloc);
    if (d.refs) (void)hipFree(d.refs);
    if (d.roles) (void)hipFree(d.roles);
    d = TableDev {};
    t = HostTable {};
}

// Arena growth: allocate the larger arena, copy the used slabs (ids are
// preserved, so every handle stays valid), release the old arena.
int table_grow(HostTable& t, uint32_t newCap, hipStream_t stream)
{
    if (newCap <= t.slabCap)
        return SVDAG_OK;
    if (((uint64_t)newCap << kSlotBits) > (1ull << 32)) {
        set_error("table_grow: slab capacity exceeds the pointer range");
        return SVDAG_E_OOM;
    }
    TableDev& d = t.dev;
    const uint32_t old = t.slabCap;
    uint32_t *lines = nullptr, *freeList = nullptr, *refs = nullptr;
    uint8_t* roles = nullptr;
    SVDAG_HIP(hipMalloc(&lines, (size_t)newCap * kLineWords * 4));
    SVDAG_HIP(hipMalloc(&freeList, (size_t)newCap * 4));
    SVDAG_HIP(hipMemcpyAsync(lines, d.lines, (size_t)old * kLineWords * 4, hipMemcpyDeviceToDevice, stream));
    SVDAG_HIP(hipMemcpyAsync(freeList, d.freeList, (size_t)old * 4, hipMemcpyDeviceToDevice, stream));
    if (d.refs) {
        SVDAG_HIP(hipMalloc(&refs, (size_t)newCap * kMaxSlots * 4));
        SVDAG_HIP(hipMalloc(&roles, (size_t)newCap * kMaxSlots));
        SVDAG_HIP(hipMemcpyAsync(refs, d.refs, (size_t)old * kMaxSlots * 4, hipMemcpyDeviceToDevice, stream));
        SVDAG_HIP(hipMemcpyAsync(roles, d.roles, (size_t)old * kMaxSlots, hipMemcpyDeviceToDevice, stream));
        SVDAG_HIP(hipMemsetAsync(refs + (size_t)old * kMaxSlots, 0, (size_t)(newCap - old) * kMaxSlots * 4, stream));
        SVDAG_HIP(hipMemsetAsync(roles + (size_t)old * kMaxSlots, 0, (size_t)(newCap - old) * kMaxSlots, stream));
    }
    SVDAG_HIP(hipStreamSynchronize(stream));
    (void)hipFree(d.lines);
    (void)hipFree(d.freeList);
    if (d.refs) {
        (void)hipFree(d.refs);
        (void)hipFree(d.roles);
        d.refs = refs;
        d.roles = roles;
    }
    d.lines = lines;
    d.freeList = freeList;
    d.slabCap = newCap;
    t.slabCap = newCap;
    table_init_lines<<<grid_for((uint64_t)(newCap - old) * kLineWords, 256, 8192), 256, 0, stream>>>(d, old, newCap);
    SVDAG_HIP(hipGetLastError());
    // The OOM flag and the over-counted fresh watermark are reset so that the
    // next allocation continues from the first never-used slab.
    AllocState a {};
    SVDAG_HIP(hipMemcpyAsync(&a, d.alloc, sizeof(a), hipMemcpyDeviceToHost, stream));
    SVDAG_HIP(hipStreamSynchronize(stream));
    a.fresh = std::min(a.fresh, old);
    a.oom = 0;
    SVDAG_HIP(hipMemcpyAsync(d.alloc, &a, sizeof(a), hipMemcpyHostToDevice, stream));
    return SVDAG_OK;
}

int table_read_alloc(const HostTable& t, AllocState& out, hipStream_t stream)
{
    SVDAG_HIP(hipMemcpyAsync(&out, t.dev.alloc, sizeof(out), hipMemcpyDeviceToHost, stream));
    SVDAG_HIP(hipStreamSynchronize(stream));
    return SVDAG_OK;
}

int TableSet::upload(hipStream_t stream)
{
    std::vector<TableDev> h(tables.size());
    for (size_t i = 0; i < tables.size(); ++i)
        h[i] = tables[i].dev;
    if (!devTables)
        SVDAG_HIP(hipMalloc(&devTables, sizeof(TableDev) * std::max<size_t>(h.size(), 1)));
    SVDAG_HIP(hipMemcpyAsync(devTables, h.data(), sizeof(TableDev) * h.size(), hipMemcpyHostToDevice, stream));
    return SVDAG_OK;
}

void TableSet::release()
{
    for (auto& t : tables)
        table_free(t);
    tables.clear();
    if (devTables)
        (void)hipFree(devTables);
    devTables = nullptr;
}

// ------------------------------------------------------------------ bulk kernels
// Segment descriptors live in device memory as 5 words each:
// [item_words, count, word_offset, out_offset, first_global_index].
__device__ inline uint32_t seg_of(const uint32_t* segs, uint32_t nseg, uint32_t idx)
{
    uint32_t s = 0;
    while (s + 1 < nseg && idx >= segs[(s + 1) * 5 + 4])
        ++s;
    return s;
}

template <int G, bool kInsert, int MW>
__global__ void __launch_bounds__(256) bulk_group_kernel(const TableDev* tables, uint32_t minWords, const uint32_t* words,
                                                         const uint32_t* segs, uint32_t nseg, uint32_t total, uint32_t* out)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const uint32_t groupsPerBlock = blockDim.x / G;
    for (uint32_t idx = blockIdx.x * groupsPerBlock + threadIdx.x / G; idx < total; idx += gridDim.x * groupsPerBlock) {
        const uint32_t sg = seg_of(segs, nseg, idx);
        const uint32_t s = segs[sg * 5 + 0];
        const uint32_t local = idx - segs[sg * 5 + 4];
        const TableDev t = tables[s - minWords];
        Needle<G, MW> n;
        needle_load(n, words + segs[sg * 5 + 2] + (size_t)local * s, s, gl);
        const uint32_t h = needle_hash(n, s, gl);
        uint32_t r;
        if constexpr (kInsert)
            r = table_insert_group<G>(t, n, s, h, gl, gbase);
        else
            r = table_find_group<G>(t, n, s, h, gl, gbase);
        if (gl == 0)
            out[segs[sg * 5 + 3] + local] = r;
    }
}

// Find with N needles per group in flight: the needle loads, hashes and tag
// line loads of all N are issued before any candidate is compared, so one
// group keeps N independent probes outstanding (memory-level parallelism
// without more waves).
template <int G, int N, int MW>
__global__ void __launch_bounds__(256) bulk_find_multi_kernel(const TableDev* tables, uint32_t minWords,
                                                              const uint32_t* words, const uint32_t* segs, uint32_t nseg,
                                                              uint32_t total, uint32_t* out)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const uint32_t groupsPerBlock = blockDim.x / G;
    const uint32_t stride = gridDim.x * groupsPerBlock * N;
    for (uint32_t base = (blockIdx.x * groupsPerBlock + threadIdx.x / G) * N; base < total; base += stride) {
        Needle<G, MW> n[N];
        uint32_t s[N], h[N], slab[N], outIdx[N];
        const TableDev* tp[N];
        bool valid[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const uint32_t idx = base + k;
            valid[k] = idx < total;
            const uint32_t sg = seg_of(segs, nseg, valid[k] ? idx : base);
            s[k] = segs[sg * 5 + 0];
            const uint32_t local = (valid[k] ? idx : base) - segs[sg * 5 + 4];
            outIdx[k] = segs[sg * 5 + 3] + local;
            tp[k] = tables + (s[k] - minWords);
            needle_load(n[k], words + segs[sg * 5 + 2] + (size_t)local * s[k], s[k], gl);
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            h[k] = needle_hash(n[k], s[k], gl);
            slab[k] = h[k] % tp[k]->numBuckets;
        }
        LineHead lines[N];
#pragma unroll
        for (int k = 0; k < N; ++k)
            line_head_load(lines[k], tp[k]->lines + (size_t)slab[k] * kLineWords, tp[k]->tagWords);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const uint32_t r = probe_chain<G, MW>(*tp[k], n[k], s[k], make_tag(h[k]) * 0x01010101u, slab[k], lines[k],
                                                  gl, gbase);
            if (gl == 0 && valid[k])
                out[outIdx[k]] = r;
        }
    }
}

template <bool kInsert>
__global__ void __launch_bounds__(256) bulk_lane_kernel(const TableDev* tables, uint32_t minWords, const uint32_t* words,
                                                        const uint32_t* segs, uint32_t nseg, uint32_t total, uint32_t* out)
{
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const uint32_t sg = seg_of(segs, nseg, idx);
        const uint32_t s = segs[sg * 5 + 0];
        const uint32_t local = idx - segs[sg * 5 + 4];
        const TableDev t = tables[s - minWords];
        const uint32_t* p = words + segs[sg * 5 + 2] + (size_t)local * s;
        uint32_t h = 0;
        for (uint32_t j = 0; j < s; ++j)
            h ^= hash_part(p[j], j);
        out[segs[sg * 5 + 3] + local] = kInsert ? table_insert_lane(t, p, s, h) : table_find_lane(t, p, s, h);
    }
}

// ------------------------------------------------------------------ XCD-binned find
// MI355X has 8 XCDs with a 4 MB L2 each, and consecutive workgroups are
// dispatched to the XCDs round robin.  A plain find probes tag lines of the
// whole table from every XCD, so no L2 holds them (22 MB of tag lines at
// 2^24 items) and every probe fetches 128 bytes from the fabric.  The binned
// find first sorts the needles by bucket range -- bucket b belongs to XCD
// b * 8 / numBuckets -- into per-(segment, XCD) bins of [out index, hash,
// words...] records, then runs blocks b with b % 8 == x over bin x only, so
// each XCD's L2 sees one eighth of the tag lines.  The find itself is the
// same table_find_group, so results are identical to the other variants.
constexpr uint32_t kXcds = 8;

struct BinDesc {
    uint32_t offset;   // first word of the bin in the bin buffer
    uint32_t cap;      // records the bin can hold
};

// Bin counters sit 128 bytes apart so the few dozen of them do not share
// L2 channels.
constexpr uint32_t kBinPad = 32;

// Four needles per thread per round; the block counts its needles per
// (segment, XCD) key in LDS, reserves each key's run with one global atomic,
// and writes the records.
template <int MW>
__global__ void __launch_bounds__(256) bin_needles(const TableDev* tables, uint32_t minWords, const uint32_t* words,
                                                   const uint32_t* segs, uint32_t nseg, uint32_t total,
                                                   const BinDesc* bins, uint32_t* binCount, uint32_t* binBuf,
                                                   uint32_t* overflow)
{
    constexpr uint32_t kPer = 4;
    __shared__ uint32_t lcount[64 * kXcds];
    __shared__ uint32_t lbase[64 * kXcds];
    const uint32_t nkeys = nseg * kXcds;
    const uint32_t chunk = blockDim.x * kPer;
    for (uint32_t c0 = blockIdx.x * chunk; c0 < total; c0 += gridDim.x * chunk) {
        for (uint32_t k = threadIdx.x; k < nkeys; k += blockDim.x)
            lcount[k] = 0u;
        __syncthreads();
        uint32_t key[kPer], local[kPer], sz[kPer], outIdx[kPer], h[kPer];
        uint32_t w[kPer][MW];
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t idx = c0 + q * blockDim.x + threadIdx.x;
            key[q] = 0xFFFFFFFFu;
            sz[q] = 0;
            h[q] = 0;
            outIdx[q] = 0;
            if (idx < total) {
                const uint32_t sg = seg_of(segs, nseg, idx);
                const uint32_t s = segs[sg * 5 + 0];
                const uint32_t lidx = idx - segs[sg * 5 + 4];
                const uint32_t* p = words + segs[sg * 5 + 2] + (size_t)lidx * s;
                uint32_t hh = 0;
#pragma unroll
                for (int j = 0; j < MW; ++j) {
                    w[q][j] = (uint32_t)j < s ? p[j] : 0u;
                    if ((uint32_t)j < s)
                        hh ^= hash_part(w[q][j], (uint32_t)j);
                }
                const uint32_t nb = tables[s - minWords].numBuckets;
                const uint32_t x = (uint32_t)(((uint64_t)(hh % nb) * kXcds) / nb);
                key[q] = sg * kXcds + x;
                sz[q] = s;
                h[q] = hh;
                outIdx[q] = segs[sg * 5 + 3] + lidx;
                local[q] = atomicAdd(&lcount[key[q]], 1u);
            }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nkeys; k += blockDim.x)
            lbase[k] = lcount[k] ? atomicAdd(&binCount[k * kBinPad], lcount[k]) : 0u;
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            if (key[q] == 0xFFFFFFFFu)
                continue;
            const uint32_t pos = lbase[key[q]] + local[q];
            const BinDesc b = bins[key[q]];
            if (pos < b.cap) {
                uint32_t* r = binBuf + b.offset + (size_t)pos * (sz[q] + 2);
                r[0] = outIdx[q];
                r[1] = h[q];
#pragma unroll
                for (int j = 0; j < MW; ++j)
                    if ((uint32_t)j < sz[q])
                        r[2 + j] = w[q][j];
            } else {
                atomicOr(overflow, 1u);
            }
        }
        __syncthreads(); // lcount is cleared next round
    }
}

// Blocks b serve XCD b % 8; N records per group in flight.
template <int G, int N, int MW>
__global__ void __launch_bounds__(256) find_binned(const TableDev* tables, uint32_t minWords, const uint32_t* segs,
                                                   uint32_t nseg, const BinDesc* bins, const uint32_t* binCount,
                                                   const uint32_t* binBuf, uint32_t* out)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const uint32_t groupsPerBlock = blockDim.x / G;
    const uint32_t x = blockIdx.x % kXcds;
    const uint32_t bx = blockIdx.x / kXcds;
    const uint32_t blocksPerXcd = gridDim.x / kXcds;
    for (uint32_t sg = 0; sg < nseg; ++sg) {
        const uint32_t s = segs[sg * 5 + 0];
        const uint32_t key = sg * kXcds + x;
        const uint32_t cnt = min(binCount[key * kBinPad], bins[key].cap);
        const uint32_t* rec0 = binBuf + bins[key].offset;
        const TableDev t = tables[s - minWords];
        const uint32_t stride = blocksPerXcd * groupsPerBlock * N;
        for (uint32_t base = (bx * groupsPerBlock + threadIdx.x / G) * N; base < cnt; base += stride) {
            Needle<G, MW> n[N];
            uint32_t h[N], slab[N], outIdx[N];
            bool valid[N];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const uint32_t e = base + k;
                valid[k] = e < cnt;
                const uint32_t* r = rec0 + (size_t)(valid[k] ? e : base) * (s + 2);
                outIdx[k] = r[0];
                h[k] = r[1];
                needle_load(n[k], r + 2, s, gl);
                slab[k] = h[k] % t.numBuckets;
            }
            LineHead lines[N];
#pragma unroll
            for (int k = 0; k < N; ++k)
                line_head_load(lines[k], t.lines + (size_t)slab[k] * kLineWords, t.tagWords);
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const uint32_t r = probe_chain<G, MW>(t, n[k], s, make_tag(h[k]) * 0x01010101u, slab[k], lines[k], gl,
                                                      gbase);
                if (gl == 0 && valid[k])
                    out[outIdx[k]] = r;
            }
        }
    }
}

// ------------------------------------------------------------------ segment-uniform per-lane find
// Blocks are assigned to segments on the host, so a block's item size and
// table are uniform: the table descriptor comes in through scalar loads and
// the segment lookup costs no memory access per lane.  Each lane carries NPL
// finds: the needle hashes and line loads of all NPL are issued before any
// candidate is compared, so a lane keeps NPL random line reads in flight.
constexpr uint32_t kMaxLaneSegs = 32;
struct LaneSegs {
    uint32_t n;
    uint32_t minWords;
    uint32_t s[kMaxLaneSegs];
    uint32_t count[kMaxLaneSegs];
    uint32_t wordOff[kMaxLaneSegs];
    uint32_t outOff[kMaxLaneSegs];
    uint32_t blockStart[kMaxLaneSegs + 1];
};

// Chain walk of a per-lane find whose first line head is already in L and
// whose needle words are in registers: a candidate's words are loaded all
// at once (no early exit), so a compare costs one round trip to the line
// the probe just fetched.
template <int MW>
__device__ inline uint32_t probe_lane(const TableDev& t, const uint32_t (&nw)[MW], uint32_t s, uint32_t tag4,
                                      uint32_t slab, LineHead L)
{
    while (true) {
        uint32_t cand = line_match(t, L, tag4);
        while (cand) {
            const uint32_t slot = __ffs(cand) - 1u;
            cand &= cand - 1u;
            const uint32_t* item = line_item(t, slab, slot);
            uint32_t w[MW];
#pragma unroll
            for (int k = 0; k < MW; ++k)
                w[k] = (uint32_t)k < s ? item[k] : 0u;
            bool eq = true;
#pragma unroll
            for (int k = 0; k < MW; ++k)
                eq &= w[k] == nw[k];
            if (eq)
                return (slab << kSlotBits) | slot;
        }
        if (L.next >= kPendingSlab)
            return kNotFound;
        slab = L.next;
        line_head_load(L, t.lines + (size_t)slab * kLineWords, t.tagWords);
    }
}

template <int NPL, int MW>
__global__ void __launch_bounds__(256) find_lanes_seg(const TableDev* tables, const uint32_t* words, LaneSegs a,
                                                      uint32_t* out)
{
    uint32_t sg = 0;
    while (sg + 1 < a.n && blockIdx.x >= a.blockStart[sg + 1])
        ++sg;
    const uint32_t s = a.s[sg];
    const TableDev t = tables[s - a.minWords];
    const uint32_t count = a.count[sg];
    const uint32_t first = (blockIdx.x - a.blockStart[sg]) * (256u * NPL) + threadIdx.x;
    const uint32_t* base = words + a.wordOff[sg];
    uint32_t nw[NPL][MW];
    uint32_t h[NPL], slab[NPL];
    LineHead L[NPL];
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const uint32_t idx = first + (uint32_t)k * 256u;
        const uint32_t* nd = base + (size_t)(idx < count ? idx : 0u) * s;
#pragma unroll
        for (int j = 0; j < MW; ++j)
            nw[k][j] = (uint32_t)j < s ? nd[j] : 0u;
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < MW; ++j)
            if ((uint32_t)j < s)
                x ^= hash_part(nw[k][j], (uint32_t)j);
        h[k] = x;
        slab[k] = x % t.numBuckets;
    }
#pragma unroll
    for (int k = 0; k < NPL; ++k)
        line_head_load(L[k], t.lines + (size_t)slab[k] * kLineWords, t.tagWords);
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
        const uint32_t idx = first + (uint32_t)k * 256u;
        const uint32_t r = probe_lane<MW>(t, nw[k], s, make_tag(h[k]) * 0x01010101u, slab[k], L[k]);
        if (idx < count)
            out[a.outOff[sg] + idx] = r;
    }
}

// Copy the words of n items (pointers already validated on the host).
__global__ void gather_items(TableDev t, const uint32_t* ptrs, uint32_t n, uint32_t* out)
{
    const uint64_t total = (uint64_t)n * t.itemWords;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(i / t.itemWords), j = (uint32_t)(i % t.itemWords);
        out[i] = table_item_ptr(t, ptrs[k])[j];
    }
}

void BinScratch::release()
{
    for (void* p : { (void*)buf, (void*)count, desc, (void*)overflow })
        if (p)
            (void)hipFree(p);
    *this = BinScratch {};
}

static int launch_binned(const TableSet& ts, const uint32_t* words, const svdag_segment* segs, uint32_t nseg,
                         const uint32_t* segBuf, uint32_t n, uint32_t* out, int variant, hipStream_t stream,
                         BinScratch* bs, hipEvent_t ev0, hipEvent_t ev1)
{
    if (!bs) {
        set_error("binned find needs bin scratch");
        return SVDAG_E_INVALID;
    }
    // Per-(segment, XCD) capacity: mean + 8 standard deviations + 256.
    std::vector<BinDesc> desc((size_t)nseg * kXcds);
    uint64_t words_ = 0;
    uint32_t maxS = 0;
    for (uint32_t sg = 0; sg < nseg; ++sg) {
        const double mean = (double)segs[sg].count / kXcds;
        const uint32_t cap = (uint32_t)(mean + 8.0 * std::sqrt(mean) + 256.0);
        maxS = std::max(maxS, segs[sg].item_words);
        for (uint32_t x = 0; x < kXcds; ++x) {
            desc[sg * kXcds + x] = BinDesc { (uint32_t)words_, cap };
            words_ += (uint64_t)cap * (segs[sg].item_words + 2);
        }
    }
    if (words_ >= (1ull << 32)) {
        set_error("binned find: too many needle words");
        return SVDAG_E_INVALID;
    }
    if (!bs->count) {
        SVDAG_HIP(hipMalloc(&bs->count, 64 * kXcds * kBinPad * 4));
        SVDAG_HIP(hipMalloc(&bs->desc, 64 * kXcds * sizeof(BinDesc)));
        SVDAG_HIP(hipMalloc(&bs->overflow, 4));
    }
    if (bs->bufWords < words_) {
        if (bs->buf)
            (void)hipFree(bs->buf);
        bs->buf = nullptr;
        bs->bufWords = 0;
        SVDAG_HIP(hipMalloc(&bs->buf, words_ * 4));
        bs->bufWords = words_;
    }
    SVDAG_HIP(hipMemcpyAsync(bs->desc, desc.data(), desc.size() * sizeof(BinDesc), hipMemcpyHostToDevice, stream));
    SVDAG_HIP(hipMemsetAsync(bs->count, 0, (size_t)nseg * kXcds * kBinPad * 4, stream));
    SVDAG_HIP(hipMemsetAsync(bs->overflow, 0, 4, stream));
    const BinDesc* dDesc = reinterpret_cast<const BinDesc*>(bs->desc);
    const bool narrow = maxS <= 10;
    const uint32_t gridBin = grid_for(n, 256 * 4, 4096);
    if (ev0)
        SVDAG_HIP(hipEventRecord(ev0, stream));
    if (narrow)
        bin_needles<10><<<gridBin, 256, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, dDesc,
                                                     bs->count, bs->buf, bs->overflow);
    else
        bin_needles<18><<<gridBin, 256, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, dDesc,
                                                     bs->count, bs->buf, bs->overflow);
    const uint32_t grid = kXcds * 512;
#define SVDAG_BINNED_CASE(V, G, N)                                                                                   \
    case V:                                                                                                          \
        if (narrow)                                                                                                  \
            find_binned<G, N, 10><<<grid, 256, 0, stream>>>(ts.devTables, ts.minWords, segBuf, nseg, dDesc,          \
                                                            bs->count, bs->buf, out);                               \
        else                                                                                                         \
            find_binned<G, N, 18><<<grid, 256, 0, stream>>>(ts.devTables, ts.minWords, segBuf, nseg, dDesc,          \
                                                            bs->count, bs->buf, out);                               \
        break;
    switch (variant) {
        SVDAG_BINNED_CASE(1004, 4, 1)
        SVDAG_BINNED_CASE(1008, 8, 1)
        SVDAG_BINNED_CASE(1204, 4, 2)
        SVDAG_BINNED_CASE(1208, 8, 2)
    default:
        set_error("unknown binned variant");
        return SVDAG_E_INVALID;
    }
#undef SVDAG_BINNED_CASE
    if (ev1)
        SVDAG_HIP(hipEventRecord(ev1, stream));
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

int launch_bulk(const TableSet& ts,const uint32_t* words, const svdag_segment* segs, uint32_t nseg, uint32_t* out,
                int variant, bool insert, hipStream_t stream, uint32_t* segBuf, BinScratch* bins, hipEvent_t ev0,
                hipEvent_t ev1)
{
    // Nothing to launch: the events still bracket an empty interval.
    auto empty = [&]() -> int {
        if (ev0)
            SVDAG_HIP(hipEventRecord(ev0, stream));
        if (ev1)
            SVDAG_HIP(hipEventRecord(ev1, stream));
        return SVDAG_OK;
    };
    if (nseg == 0)
        return empty();
    if (nseg > 64) {
        set_error("too many segments (max 64)");
        return SVDAG_E_INVALID;
    }
    uint32_t host[64 * 5];
    uint64_t total = 0;
    for (uint32_t i = 0; i < nseg; ++i) {
        if (segs[i].item_words < ts.minWords || segs[i].item_words > ts.maxWords) {
            set_error("segment item size outside the table set");
            return SVDAG_E_INVALID;
        }
        host[i * 5 + 0] = segs[i].item_words;
        host[i * 5 + 1] = segs[i].count;
        host[i * 5 + 2] = segs[i].word_offset;
        host[i * 5 + 3] = segs[i].out_offset;
        host[i * 5 + 4] = (uint32_t)total;
        total += segs[i].count;
    }
    if (total >= (1ull << 32)) {
        set_error("too many items");
        return SVDAG_E_INVALID;
    }
    // The segment-uniform lane variants (1, 2, 3) take their segments as
    // kernel arguments; the others read them from segBuf.
    if (!(variant >= 1 && variant <= 3))
        SVDAG_HIP(hipMemcpyAsync(segBuf, host, nseg * 5 * 4, hipMemcpyHostToDevice, stream));
    const uint32_t n = (uint32_t)total;
    if (n == 0)
        return empty();
    if (variant >= 1000) {
        if (insert) {
            set_error("binned variants (1004, 1008, 1204, 1208) are find only");
            return SVDAG_E_INVALID;
        }
        return launch_binned(ts, words, segs, nseg, segBuf, n, out, variant, stream, bins, ev0, ev1);
    }
    const uint32_t block = 256;
    if (variant >= 1 && variant <= 3) {
        if (insert) {
            set_error("segment-uniform lane variants (1, 2, 3) are find only");
            return SVDAG_E_INVALID;
        }
        if (nseg > kMaxLaneSegs) {
            set_error("segment-uniform lane variants take at most 32 segments");
            return SVDAG_E_INVALID;
        }
        const int npl = variant == 1 ? 1 : (variant == 2 ? 2 : 4);
        LaneSegs a {};
        a.n = nseg;
        a.minWords = ts.minWords;
        uint64_t blocks = 0;
        for (uint32_t i = 0; i < nseg; ++i) {
            a.s[i] = segs[i].item_words;
            a.count[i] = segs[i].count;
            a.wordOff[i] = segs[i].word_offset;
            a.outOff[i] = segs[i].out_offset;
            a.blockStart[i] = (uint32_t)blocks;
            blocks += ((uint64_t)segs[i].count + 256u * npl - 1) / (256u * npl);
        }
        a.blockStart[nseg] = (uint32_t)blocks;
        if (ev0)
            SVDAG_HIP(hipEventRecord(ev0, stream));
        const bool narrowItems = ts.maxWords <= 10;
#define SVDAG_LANE_SEG(NPL)                                                                                  \
        if (narrowItems)                                                                                     \
            find_lanes_seg<NPL, 10><<<(uint32_t)blocks, block, 0, stream>>>(ts.devTables, words, a, out);    \
        else                                                                                                 \
            find_lanes_seg<NPL, 18><<<(uint32_t)blocks, block, 0, stream>>>(ts.devTables, words, a, out);
        if (npl == 1) {
            SVDAG_LANE_SEG(1)
        } else if (npl == 2) {
            SVDAG_LANE_SEG(2)
        } else {
            SVDAG_LANE_SEG(4)
        }
#undef SVDAG_LANE_SEG
        if (ev1)
            SVDAG_HIP(hipEventRecord(ev1, stream));
        SVDAG_HIP(hipGetLastError());
        return SVDAG_OK;
    }
    if (ev0)
        SVDAG_HIP(hipEventRecord(ev0, stream));
    // Items of at most 10 words need fewer needle registers per lane.
    const bool narrow = ts.maxWords <= 10;
    switch (variant) {
    case 0: {
        const uint32_t grid = grid_for(n, block, 1u << 16);
        if (insert)
            bulk_lane_kernel<true><<<grid, block, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, out);
        else
            bulk_lane_kernel<false><<<grid, block, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, out);
        break;
    }
#define SVDAG_BULK_LAUNCH(G, INS, MW) \
    bulk_group_kernel<G, INS, MW><<<grid, block, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, out)
#define SVDAG_BULK_CASE(G)                                                                                              \
    case G: {                                                                                                           \
        const uint32_t grid = grid_for((uint64_t)n * G, block, 1u << 16);                                               \
        if (insert) {                                                                                                   \
            if (narrow) SVDAG_BULK_LAUNCH(G, true, 10); else SVDAG_BULK_LAUNCH(G, true, 18);                            \
        } else {                                                                                                        \
            if (narrow) SVDAG_BULK_LAUNCH(G, false, 10); else SVDAG_BULK_LAUNCH(G, false, 18);                          \
        }                                                                                                               \
        break;                                                                                                          \
    }
        SVDAG_BULK_CASE(4)
        SVDAG_BULK_CASE(8)
        SVDAG_BULK_CASE(16)
        SVDAG_BULK_CASE(32)
#undef SVDAG_BULK_CASE
#undef SVDAG_BULK_LAUNCH
#define SVDAG_MULTI_CASE(V, G, N)                                                                                       \
    case V: {                                                                                                           \
        if (insert) {                                                                                                   \
            set_error("multi-needle variants are find-only");                                                           \
            return SVDAG_E_INVALID;                                                                                     \
        }                                                                                                               \
        const uint32_t grid = grid_for(((uint64_t)n + N - 1) / N * G, block, 1u << 16);                                 \
        if (narrow)                                                                                                     \
            bulk_find_multi_kernel<G, N, 10><<<grid, block, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, out); \
        else                                                                                                            \
            bulk_find_multi_kernel<G, N, 18><<<grid, block, 0, stream>>>(ts.devTables, ts.minWords, words, segBuf, nseg, n, out); \
        break;                                                                                                          \
    }
        SVDAG_MULTI_CASE(204, 4, 2)
        SVDAG_MULTI_CASE(404, 4, 4)
        SVDAG_MULTI_CASE(208, 8, 2)
        SVDAG_MULTI_CASE(408, 8, 4)
#undef SVDAG_MULTI_CASE
    default:
        set_error("variant must be 0, 4, 8, 16, 32 (lanes per item), 1, 2, 3 (lanes, 1/2/4 finds each, find only), "
                  "204, 404, 208, 408 (N*100 + G, find only) "
                  "or 1004, 1008, 1204, 1208 (XCD-binned, find only)");
        return SVDAG_E_INVALID;
    }
    if (ev1)
        SVDAG_HIP(hipEventRecord(ev1, stream));
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

} // namespace svdag

using namespace svdag;

namespace svdag {
int require_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device visible");
        return SVDAG_E_NO_DEVICE;
    }
    return SVDAG_OK;
}
} // namespace svdag

struct svdag_hashset {
    TableSet ts;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t regionStart = nullptr, regionEnd = nullptr; // svdag_hashset_timer
    uint32_t* segBuf = nullptr;
    BinScratch bins;
};

extern "C" {

const char* svdag_last_error(void) { return g_lastError.c_str(); }
uint32_t svdag_abi_version(void) { return 2u; }

int svdag_device_count(int* out)
{
    if (!out)
        return SVDAG_E_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        n = 0;
    *out = n;
    return SVDAG_OK;
}

int svdag_dev_alloc(size_t bytes, void** out)
{
    if (!out)
        return SVDAG_E_INVALID;
    SVDAG_TRY(require_device());
    SVDAG_HIP(hipMalloc(out, bytes ? bytes : 4));
    return SVDAG_OK;
}
int svdag_dev_free(void* p)
{
    if (p)
        SVDAG_HIP(hipFree(p));
    return SVDAG_OK;
}
int svdag_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    SVDAG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SVDAG_OK;
}
int svdag_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    SVDAG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SVDAG_OK;
}
int svdag_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    SVDAG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return SVDAG_OK;
}
int svdag_device_synchronize(void)
{
    SVDAG_HIP(hipDeviceSynchronize());
    return SVDAG_OK;
}

int svdag_hashset_create(uint32_t minWords, uint32_t maxWords, const uint32_t* buckets, const uint32_t* reserve,
                         svdag_hashset** out)
{
    if (!out || !buckets || minWords == 0 || maxWords < minWords || maxWords > kMaxItemWords) {
        set_error("svdag_hashset_create: bad arguments");
        return SVDAG_E_INVALID;
    }
    SVDAG_TRY(require_device());
    auto* hs = new svdag_hashset();
    SVDAG_HIP(hipStreamCreateWithFlags(&hs->stream, hipStreamNonBlocking));
    SVDAG_HIP(hipEventCreate(&hs->ev0));
    SVDAG_HIP(hipEventCreate(&hs->ev1));
    SVDAG_HIP(hipEventCreate(&hs->regionStart));
    SVDAG_HIP(hipEventCreate(&hs->regionEnd));
    SVDAG_HIP(hipMalloc(&hs->segBuf, 64 * 5 * 4));
    hs->ts.minWords = minWords;
    hs->ts.maxWords = maxWords;
    hs->ts.tables.resize(maxWords - minWords + 1);
    for (uint32_t s = minWords; s <= maxWords; ++s) {
        const uint32_t nb = std::max(1u, buckets[s - minWords]);
        const uint64_t res = reserve ? reserve[s - minWords] : 0;
        // Overflow reserve: equal-hash items (bulk multiset inserts) need a
        // line per `slots` at worst; in a table sized for 75% of its slots
        // about a fifth of the buckets overflow into a second line.
        const uint32_t slots = line_slots(s);
        uint64_t cap = (uint64_t)nb + res / slots + nb / 2 + 64;
        cap = std::min<uint64_t>(cap, (1ull << (32 - kSlotBits)));
        int r = table_alloc(hs->ts.tables[s - minWords], s, nb, (uint32_t)cap, hs->stream);
        if (r != SVDAG_OK) {
            hs->ts.release();
            delete hs;
            return r;
        }
    }
    SVDAG_TRY(hs->ts.upload(hs->stream));
    SVDAG_HIP(hipStreamSynchronize(hs->stream));
    *out = hs;
    return SVDAG_OK;
}

int svdag_hashset_destroy(svdag_hashset* hs)
{
    if (!hs)
        return SVDAG_OK;
    (void)hipStreamSynchronize(hs->stream);
    hs->ts.release();
    hs->bins.release();
    (void)hipFree(hs->segBuf);
    (void)hipEventDestroy(hs->ev0);
    (void)hipEventDestroy(hs->ev1);
    if (hs->regionStart)
        (void)hipEventDestroy(hs->regionStart);
    if (hs->regionEnd)
        (void)hipEventDestroy(hs->regionEnd);
    (void)hipStreamDestroy(hs->stream);
    delete hs;
    return SVDAG_OK;
}

static int hashset_bulk(svdag_hashset* hs, const uint32_t* words, const svdag_segment* segs, uint32_t nseg,
                        uint32_t* out, int variant, float* ms, bool insert)
{
    if (!hs || !words || !segs || !out)
        return SVDAG_E_INVALID;
    SVDAG_TRY(launch_bulk(hs->ts, words, segs, nseg, out, variant, insert, hs->stream, hs->segBuf, &hs->bins,
                          ms ? hs->ev0 : nullptr, ms ? hs->ev1 : nullptr));
    // An untimed find (not XCD-binned, which checks its bins afterwards) is
    // only queued: later calls on the set run after it on its stream.
    if (!ms && !insert && variant < 1000)
        return SVDAG_OK;
    SVDAG_HIP(hipStreamSynchronize(hs->stream));
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, hs->ev0, hs->ev1));
    if (variant >= 1000 && !insert) {
        // A bin that overflowed (far beyond 8 standard deviations) leaves
        // records unprocessed: redo the whole find with the plain group kernel.
        uint32_t ovf = 0;
        SVDAG_HIP(hipMemcpy(&ovf, hs->bins.overflow, 4, hipMemcpyDeviceToHost));
        if (ovf)
            return hashset_bulk(hs, words, segs, nseg, out, variant % 100, ms, insert);
    }
    if (insert) {
        for (auto& t : hs->ts.tables) {
            AllocState a {};
            SVDAG_TRY(table_read_alloc(t, a, hs->stream));
            if (a.oom) {
                set_error("slab arena exhausted during bulk insert");
                return SVDAG_E_OOM;
            }
        }
    }
    return SVDAG_OK;
}

int svdag_hashset_add_dev(svdag_hashset* hs, const uint32_t* words, const svdag_segment* segs, uint32_t nseg,
                          uint32_t* out, int variant, float* ms)
{
    return hashset_bulk(hs, words, segs, nseg, out, variant, ms, true);
}

int svdag_hashset_find_dev(svdag_hashset* hs, const uint32_t* words, const svdag_segment* segs, uint32_t nseg,
                           uint32_t* out, int variant, float* ms)
{
    return hashset_bulk(hs, words, segs, nseg, out, variant, ms, false);
}

int svdag_hashset_timer(svdag_hashset* hs, uint32_t stop, float* elapsed_ms)
{
    if (!hs || (stop && !elapsed_ms))
        return SVDAG_E_INVALID;
    if (!stop) {
        SVDAG_HIP(hipEventRecord(hs->regionStart, hs->stream));
        return SVDAG_OK;
    }
    SVDAG_HIP(hipEventRecord(hs->regionEnd, hs->stream));
    SVDAG_HIP(hipEventSynchronize(hs->regionEnd));
    SVDAG_HIP(hipEventElapsedTime(elapsed_ms, hs->regionStart, hs->regionEnd));
    return SVDAG_OK;
}

int svdag_hashset_read(svdag_hashset* hs, uint32_t s, const uint32_t* ptrs, uint32_t n, uint32_t* outWords)
{
    if (!hs || s < hs->ts.minWords || s > hs->ts.maxWords || (n && (!ptrs || !outWords)))
        return SVDAG_E_INVALID;
    if (n == 0)
        return SVDAG_OK;
    const HostTable& t = hs->ts.tables[s - hs->ts.minWords];
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t slab = ptrs[i] >> kSlotBits, slot = ptrs[i] & ((1u << kSlotBits) - 1u);
        if (ptrs[i] == kNotFound || slab >= t.slabCap || slot >= t.dev.slots) {
            set_error("svdag_hashset_read: invalid pointer");
            return SVDAG_E_INVALID;
        }
    }
    uint32_t *dPtrs = nullptr, *dOut = nullptr;
    SVDAG_HIP(hipMalloc(&dPtrs, (size_t)n * 4));
    SVDAG_HIP(hipMalloc(&dOut, (size_t)n * s * 4));
    SVDAG_HIP(hipMemcpyAsync(dPtrs, ptrs, (size_t)n * 4, hipMemcpyHostToDevice, hs->stream));
    gather_items<<<grid_for((uint64_t)n * s, 256, 1u << 16), 256, 0, hs->stream>>>(t.dev, dPtrs, n, dOut);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipMemcpyAsync(outWords, dOut, (size_t)n * s * 4, hipMemcpyDeviceToHost, hs->stream));
    SVDAG_HIP(hipStreamSynchronize(hs->stream));
    (void)hipFree(dPtrs);
    (void)hipFree(dOut);
    return SVDAG_OK;
}

int svdag_hashset_usage(svdag_hashset* hs, uint32_t s, uint32_t out[4])
{
    if (!hs || !out || s < hs->ts.minWords || s > hs->ts.maxWords)
        return SVDAG_E_INVALID;
    const HostTable& t = hs->ts.tables[s - hs->ts.minWords];
    AllocState a {};
    SVDAG_TRY(table_read_alloc(t, a, hs->stream));
    out[0] = t.numBuckets;
    out[1] = t.slabCap;
    out[2] = std::min(a.fresh, t.slabCap); // slabs ever handed out (head + overflow)
    out[3] = a.oom;
    return SVDAG_OK;
}

} // extern "C"
