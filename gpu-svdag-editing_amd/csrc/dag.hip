// The editable DAG: context, fully-filled chain, the edit pipeline (top-down
// edit-SVO construction, bottom-up merge into the hash-consed DAG), terrain
// build, voxel queries and export.
//
// reference: MyGPUHashDAG (my_gpu_hash_dag.{h,cu}), create_edit_svo.cu,
// my_gpu_hash_dag_edits.cu, MyGPUHashDAGUndoRedo (my_gpu_base_dag.h:16-82).
//
// Execution model: every kernel reads its item count from device memory and
// walks it with a grid-stride loop, so one edit is a fixed sequence of
// launches on one stream with no host round trip until the final status
// read.  Work items of all levels live in one array (level L+1 starts where
// level L ends); the SVO node of work item w is nodes[w], the SVO leaf of
// leaf work item i is leaves[i].
#include "dag_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>

namespace svdag {

// Full-item "counts" only gate discovery on the level above (nonzero or
// not), so they are flags: a cached load first keeps thousands of full
// items from serialising on one atomic.
__device__ inline void set_flag(uint32_t* p)
{
    if (ld_filter(p) == 0u)
        atomicOr(p, 1u);
}

__device__ inline uint32_t dedup_cap(uint32_t count, uint32_t maxCap)
{
    uint32_t cap = 64;
    while (cap < 2u * count && cap < maxCap)
        cap <<= 1;
    return cap;
}

// The reference drops an edit whose SVO has no leaf (my_gpu_hash_dag_edits.cu:651).
// A build is not an edit: a grid or terrain made only of fully-filled nodes
// has no leaf either, and its merged root is the DAG.
__device__ inline bool merge_skipped(const WsDev& ws, const DagDev& dag)
{
    return ws.status->overflow != 0 || (ws.lvl[dag.leafLevel].count == 0 && !ws.status->build);
}

// ---------------------------------------------------------------- fully-filled chain
// my_gpu_hash_dag.cu:38-63: a full leaf per material, then per level a node
// with 8 copies of the level below, flagged full.
__global__ void init_fully_filled(DagDev dag)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    uint32_t item[kMaxItemWords];
    for (uint32_t m = 0; m < dag.numMaterials; ++m) {
        item[0] = 0xFFFFFFFFu;
        item[1] = 0xFFFFFFFFu;
        const uint32_t matWords = (64u * dag.M) / 32u;
        const uint32_t rep = dag.M == 0 ? 0u : (dag.M == 4 ? m * 0x11111111u : m * 0x01010101u);
        for (uint32_t i = 0; i < matWords; ++i)
            item[2 + i] = rep;
        const uint32_t s = 2 + matWords;
        uint32_t h = 0;
        for (uint32_t j = 0; j < s; ++j)
            h ^= hash_part(item[j], j);
        const TableDev& t = dag.tables[s - 2];
        uint32_t r = table_find_lane(t, item, s, h);
        if (r == kNotFound) {
            r = table_insert_lane(t, item, s, h);
            t.roles[r] = 0u;
            ++*dag.stored;
        }
        dag.fullyFilled[dag.leafLevel * dag.numMaterials + m] = leaf_handle(dag, r, s);
    }
    for (int level = (int)dag.leafLevel - 1; level >= 0; --level) {
        for (uint32_t m = 0; m < dag.numMaterials; ++m) {
            item[0] = header_make_full(0xFFu, m, dag.M);
            for (uint32_t c = 0; c < 8; ++c)
                item[1 + c] = dag.fullyFilled[(level + 1) * dag.numMaterials + m];
            uint32_t h = 0;
            for (uint32_t j = 0; j < 9; ++j)
                h ^= hash_part(item[j], j);
            const TableDev& t = dag.tables[9 - 2];
            uint32_t r = table_find_lane(t, item, 9, h);
            if (r == kNotFound) {
                r = table_insert_lane(t, item, 9, h);
                const uint8_t role = node_role(dag, (uint32_t)level);
                t.roles[r] = role;
                for (uint32_t c = 0; c < 8; ++c)
                    child_ref_add(dag, item[1 + c], role);
                ++*dag.stored;
            }
            // The level-0 nodes are collection roots for good: one
            // permanent reference keeps their chains stored.
            if (level == 0)
                atomicAdd(&t.refs[r], 1u);
            dag.fullyFilled[level * dag.numMaterials + m] = node_handle(r, 9);
        }
    }
}

// ---------------------------------------------------------------- edit kernels
__global__ void edit_begin(WsDev ws, const uint32_t* root, uint32_t isCopy, uint32_t isBuild)
{
    const uint32_t t = threadIdx.x;
    if (t < kMaxLevels) {
        LevelDesc d {};
        if (t == 0)
            d.count = 1;
        ws.lvl[t] = d;
    }
    if (t == 0) {
        const uint32_t r = *root;
        EditStatus st {};
        st.rootBefore = r;
        st.rootAfter = r;
        st.build = isBuild;
        *ws.status = st;
        ws.work[0] = make_uint4(0u, 0u, 0u, r);
        if (isCopy) {
            CopyState cs;
            for (int k = 0; k < 8; ++k)
                cs.p[k] = CopyParent { 0u, 0u, 0u, kSentinel };
            cs.p[0].h = r;
            ws.state[0] = cs;
        }
    }
}

// One work item per 8 lanes, one lane per child (create_edit_svo.cu:167-303).
// kMulti: items of several edits share the level; item w uses eds[editOf[w]]
// and its children inherit the edit.
//
// Child slots are reserved once per block and iteration: waves count their
// subdivided children with a ballot, the block sums them in LDS and one lane
// does the single atomicAdd, so the level counter is not a hot spot.  The
// per-edit counts of a fused wave go through an LDS histogram (the edits of
// a wave are at most 256 consecutive indices, so e & 255 is a unique key)
// flushed once per block.
template <bool kMulti>
__global__ void __launch_bounds__(256) svo_inner_level(WsDev ws, DagDev dag, const EditorDev* eds, uint32_t L)
{
    __shared__ uint32_t waveTotal[4];
    __shared__ uint32_t blockBase;
    __shared__ uint32_t hist[kMulti ? 256 : 1];
    __shared__ uint32_t histEdit[kMulti ? 256 : 1];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t c = lane & 7u;
    const uint32_t gbase = lane & ~7u;
    const LevelDesc me = ws.lvl[L];
    const uint32_t childStart = me.start + me.count;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        ws.lvl[L + 1].start = childStart;
    if (kMulti) {
        hist[threadIdx.x] = 0u;
        __syncthreads();
    }
    const uint32_t childLevel = L + 1;
    const uint32_t childDepth = dag.levels - childLevel;
    const bool leafChildren = childLevel == dag.leafLevel;
    const uint32_t groupsPerBlock = blockDim.x / 8u;
    // After an overflow the count may run past the array; the edit is
    // discarded then, but no lane may read or write beyond capWork.
    const uint32_t count = me.start >= ws.capWork ? 0u : min(me.count, ws.capWork - me.start);
    for (uint32_t i0 = blockIdx.x * groupsPerBlock; i0 < count; i0 += gridDim.x * groupsPerBlock) {
        const uint32_t i = i0 + threadIdx.x / 8u;
        const bool active = i < count;
        const uint32_t w = me.start + (active ? i : 0u);
        const uint32_t e = (kMulti && active) ? ws.editOf[w] : 0u;
        const EditorDev& ed = eds[e];
        const uint4 item = active ? ws.work[w] : make_uint4(0u, 0u, 0u, kSentinel);
        uint32_t childHandle = kSentinel;
        if (item.w != kSentinel) {
            const uint32_t* node = dag_node_ptr(dag, item.w);
            const uint32_t mask = node[0] & 0xFFu;
            if ((mask >> c) & 1u)
                childHandle = node[1 + __builtin_popcount(mask & ((1u << c) - 1u))];
        }
        const uint32_t cx = (item.x << 1) | ((c >> 2) & 1u);
        const uint32_t cy = (item.y << 1) | ((c >> 1) & 1u);
        const uint32_t cz = (item.z << 1) | (c & 1u);
        NState st = NS_PARTIAL;
        uint32_t stMat = 0;
        // Only the paint tool's decision reads Filled vs Partial; the others
        // skip the child header read.
        if (childHandle == kSentinel)
            st = NS_EMPTY;
        else if (!leafChildren && ed.kind == EK_PAINT)
            st = header_full(dag_node_ptr(dag, childHandle)[0], dag.M, stMat) ? NS_FILLED : NS_PARTIAL;
        uint32_t fillMat = 0;
        Op op = OP_KEEP;
        CopyState cs;
        if (active) {
            if (ed.kind == EK_COPY)
                op = decide_copy(ed, dag, cx, cy, cz, childDepth, ws.state[w], cs);
            else
                op = decide(ed, cx, cy, cz, childDepth, st, fillMat);
        }
        const bool sub = active && op == OP_SUBDIVIDE;
        const unsigned long long subBal = __ballot(sub);
        const uint32_t subMask = (uint32_t)(subBal >> gbase) & 0xFFu;
        // Block-wide reservation of the children.
        if (lane == 0)
            waveTotal[wv] = (uint32_t)__builtin_popcountll(subBal);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t k = 0; k < blockDim.x / 64u; ++k)
                tot += waveTotal[k];
            blockBase = tot ? atomicAdd(&ws.lvl[L + 1].count, tot) : 0u;
        }
        __syncthreads();
        uint32_t base = blockBase;
        for (uint32_t k = 0; k < wv; ++k)
            base += waveTotal[k];
        if (kMulti && c == 0 && subMask) {
            atomicAdd(&hist[e & 255u], (uint32_t)__builtin_popcount(subMask));
            histEdit[e & 255u] = e;
        }
        if (sub) {
            const uint32_t local = base + (uint32_t)__builtin_popcountll(subBal & ((1ull << lane) - 1ull));
            const uint64_t g = (uint64_t)childStart + local;
            if (g < ws.capWork && (!leafChildren || local < ws.capLeaf)) {
                ws.work[g] = make_uint4(cx, cy, cz, childHandle);
                if (kMulti)
                    ws.editOf[g] = e;
                if (ed.kind == EK_COPY)
                    ws.state[g] = cs;
            } else {
                atomicOr(&ws.status->overflow, 1u);
            }
            childHandle = local;
        } else if (op == OP_EMPTY) {
            childHandle = kSentinel;
        } else if (op == OP_FILL) {
            childHandle = dag.fullyFilled[childLevel * dag.numMaterials + fillMat];
        }
        const bool out = active && childHandle != kSentinel;
        const uint32_t cmask = group_ballot<8>(out, gbase);
        if (active) {
            uint32_t* rec = ws.nodes + (size_t)w * kNodeStride;
            const uint32_t n = __builtin_popcount(cmask);
            if (out)
                rec[1 + __builtin_popcount(cmask & ((1u << c) - 1u))] = childHandle;
            if (c >= n)
                rec[1 + c] = 0u;
            if (c == 0) {
                rec[0] = cmask ? header_make(cmask, subMask, dag.M) : 0u;
                rec[9] = 0u;
            }
        }
        __syncthreads(); // waveTotal / blockBase are rewritten next iteration
    }
    if (kMulti) {
        __syncthreads();
        const uint32_t v = hist[threadIdx.x];
        if (v)
            atomicAdd(&ws.edCount[histEdit[threadIdx.x] * kMaxLevels + L + 1], v);
    }
}
// kLeavesPerWave leaves per wave64: 64 / kLeavesPerWave lanes per leaf, lane
// s deciding voxels s, s + kLeafLanes, ... of it (create_edit_svo.cu:320-426).
// A leaf's inputs (work item, editor, old leaf) are a chain of dependent
// loads, so a wave keeps several such chains in flight instead of one.
#ifndef SVDAG_LEAVES_PER_WAVE
#define SVDAG_LEAVES_PER_WAVE 4
#endif
constexpr uint32_t kLeavesPerWave = SVDAG_LEAVES_PER_WAVE;
constexpr uint32_t kLeafLanes = 64u / kLeavesPerWave; // lanes per leaf
constexpr uint32_t kLeafWordsPerLane = (18u + kLeafLanes - 1u) / kLeafLanes; // a record has at most 18 words

template <bool kMulti>
__global__ void __launch_bounds__(256) svo_leaf_level(WsDev ws, DagDev dag, const EditorDev* eds)
{
    __shared__ uint32_t lds[4][kLeavesPerWave][16];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t seg = lane / kLeafLanes, s = lane % kLeafLanes;
    const LevelDesc me = ws.lvl[dag.leafLevel];
    const uint32_t wavesPerBlock = blockDim.x / 64u;
    const uint32_t count = min(me.count, ws.capLeaf);
    for (uint32_t i0 = (blockIdx.x * wavesPerBlock + wv) * kLeavesPerWave; i0 < count;
         i0 += gridDim.x * wavesPerBlock * kLeavesPerWave) {
        const uint32_t i = i0 + seg;
        const uint32_t w = me.start + i;
        const bool valid = i < count && w < ws.capWork;
        uint32_t has4 = 0, mats[kLeavesPerWave];
#pragma unroll
        for (uint32_t v = 0; v < kLeavesPerWave; ++v)
            mats[v] = 0u;
        uint32_t oldH = kSentinel;
        const uint32_t* oldLeaf = nullptr;
        if (valid) {
            const EditorDev& ed = eds[kMulti ? ws.editOf[w] : 0u];
            const uint4 item = ws.work[w];
            oldH = item.w;
            uint64_t oldBits = 0;
            if (oldH != kSentinel) {
                oldLeaf = dag_leaf_ptr(dag, oldH);
                oldBits = (uint64_t)oldLeaf[0] | ((uint64_t)oldLeaf[1] << 32);
            }
#pragma unroll
            for (uint32_t v = 0; v < kLeavesPerWave; ++v) {
                const uint32_t b = v * kLeafLanes + s;
                const bool hasOld = (oldBits >> b) & 1ull;
                uint32_t oldMat = 0;
                if (hasOld && dag.M) {
                    const uint32_t off = (uint32_t)__builtin_popcountll(oldBits & ((1ull << b) - 1ull)) * dag.M;
                    oldMat = (oldLeaf[2 + (off >> 5)] >> (off & 31u)) & ((1u << dag.M) - 1u);
                }
                const uint32_t vx = (item.x << 2) | (((b >> 5) & 1u) << 1) | ((b >> 2) & 1u);
                const uint32_t vy = (item.y << 2) | (((b >> 4) & 1u) << 1) | ((b >> 1) & 1u);
                const uint32_t vz = (item.z << 2) | (((b >> 3) & 1u) << 1) | (b & 1u);
                uint32_t fm;
                const Op op = decide(ed, vx, vy, vz, 0, hasOld ? NS_FILLED : NS_EMPTY, fm);
                bool has = false;
                uint32_t mat = 0;
                if (ed.kind == EK_DENSE || ed.kind == EK_TERRAIN) {
                    // a set voxel of the grid or below the surface, with its own material;
                    // a tiled terrain keeps voxels outside its tile
                    has = op == OP_FILL || (op == OP_KEEP && hasOld);
                    mat = op == OP_FILL ? fm : oldMat;
                } else if (op == OP_FILL || op == OP_SUBDIVIDE) {
                    if (ed.kind == EK_COPY)
                        has = copy_value(ed, dag, vx, vy, vz, ws.state[w], hasOld, oldMat, mat);
                    else
                        has = new_value(ed, hasOld, oldMat, mat);
                } else if (op == OP_KEEP) {
                    has = hasOld;
                    mat = oldMat;
                }
                has4 |= (has ? 1u : 0u) << v;
                mats[v] = mat;
            }
        }
        // The leaf's 64-bit mask: voxel v * 16 + s is bit s of this leaf's
        // sixteen lanes in the ballot of voxel slot v.
        uint64_t mask = 0;
#pragma unroll
        for (uint32_t v = 0; v < kLeavesPerWave; ++v) {
            const uint64_t bal = __ballot((has4 >> v) & 1u);
            mask |= ((bal >> (kLeafLanes * seg)) & ((1ull << kLeafLanes) - 1ull)) << (kLeafLanes * v);
        }
        for (uint32_t q = s; q < 16u; q += kLeafLanes)
            lds[wv][seg][q] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (dag.M) {
#pragma unroll
            for (uint32_t v = 0; v < kLeavesPerWave; ++v) {
                if ((has4 >> v) & 1u) {
                    const uint32_t b = v * kLeafLanes + s;
                    const uint32_t off = (uint32_t)__builtin_popcountll(mask & ((1ull << b) - 1ull)) * dag.M;
                    atomicOr(&lds[wv][seg][off >> 5], mats[v] << (off & 31u));
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // Words s, s + kLeafLanes, ... of the record.
        uint32_t word[kLeafWordsPerLane];
#pragma unroll
        for (uint32_t k = 0; k < kLeafWordsPerLane; ++k) {
            const uint32_t q = s + kLeafLanes * k;
            word[k] = q == 0u ? (uint32_t)mask : q == 1u ? (uint32_t)(mask >> 32) : (q - 2u < 16u ? lds[wv][seg][q - 2u] : 0u);
            if (valid && q < ws.leafStride)
                ws.leaves[(size_t)i * ws.leafStride + q] = word[k];
        }
        // A leaf the edit left as it was resolves to its old handle: no dedup
        // and no DAG find (created = 2 marks it for the merge).
        bool cmp = false, differ = false;
        if (valid && oldH != kSentinel && mask != 0ull) {
            const uint32_t newSize = leaf_size_from_words((uint32_t)mask, (uint32_t)(mask >> 32), dag.M);
            if (leaf_handle_size(dag, oldH) == newSize) {
                cmp = true;
#pragma unroll
                for (uint32_t k = 0; k < kLeafWordsPerLane; ++k) {
                    const uint32_t q = s + kLeafLanes * k;
                    if (q < newSize && oldLeaf[q] != word[k])
                        differ = true;
                }
            }
        }
        const uint64_t diff = __ballot(differ);
        const bool unchanged = cmp && ((diff >> (kLeafLanes * seg)) & ((1ull << kLeafLanes) - 1ull)) == 0ull;
        if (valid && s == 0u) {
            ws.created[w] = unchanged ? 2u : 0u;
            if (unchanged) {
                ws.rep[w] = w;
                ws.handle[w] = oldH;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// Bottom-up, per level: clear the level's dedup table; map SVO children to
// DAG handles and drop empty ones (updateChildPointers,
// my_gpu_hash_dag_edits.cu:44-95); flag full items when the level below had
// any (discover_fully_filled, 96-150, gated at 337).  kMulti: the gate and
// the full-item counts are per edit, and items of an edit that made no leaf
// are emptied (that edit is dropped as a whole, :651).
template <bool kMulti>
__global__ void __launch_bounds__(256) merge_prepare(WsDev ws, DagDev dag, uint32_t L)
{
    if (merge_skipped(ws, dag))
        return;
    const LevelDesc me = ws.lvl[L];
    const uint32_t cap = dedup_cap(me.count, ws.dedupCap);
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < cap; i += stride)
        ws.dedup[i] = 0ull;
    if (L == dag.leafLevel) {
        for (uint32_t i = tid; i < me.count; i += stride) {
            uint32_t m;
            if (leaf_is_full(ws.leaves + (size_t)i * ws.leafStride, dag.M, m))
                set_flag(kMulti ? &ws.edFull[ws.editOf[me.start + i] * kMaxLevels + L] : &ws.lvl[L].fullCount);
        }
        return;
    }
    const bool gateAll = ws.lvl[L + 1].fullCount != 0;
    const uint32_t childStart = ws.lvl[L + 1].start;
    const bool leafChildren = L + 1 == dag.leafLevel;
    for (uint32_t i = tid; i < me.count; i += stride) {
        uint32_t* rec = ws.nodes + (size_t)(me.start + i) * kNodeStride;
        const uint32_t e = kMulti ? ws.editOf[me.start + i] : 0u;
        const bool gate = kMulti ? ws.edFull[e * kMaxLevels + L + 1] != 0 : gateAll;
        const uint32_t hdr = rec[0];
        const uint32_t mask = hdr & 0xFFu;
        if (mask == 0)
            continue;
        if (kMulti && ws.edCount[e * kMaxLevels + dag.leafLevel] == 0) {
            rec[0] = 0u;
            continue;
        }
        const uint32_t imask = header_payload(hdr, dag.M) & 0xFFu;
        uint32_t out[8];
        uint32_t n = 0, newMask = 0, k = 0, newChild = 0;
        for (uint32_t c = 0; c < 8; ++c) {
            if (!((mask >> c) & 1u))
                continue;
            uint32_t h = rec[1 + k++];
            if ((imask >> c) & 1u) {
                const uint32_t r = ws.rep[childStart + h];
                h = (r == kSentinel) ? kSentinel : ws.handle[r];
                if (r != kSentinel)
                    newChild |= ws.created[r] & 1u; // 1: inserted by this merge (2: unchanged leaf)
            }
            if (h != kSentinel) {
                out[n++] = h;
                newMask |= 1u << c;
            }
        }
        bool full = false;
        uint32_t fm = 0;
        if (gate && newMask == 0xFFu) {
            full = true;
            for (uint32_t j = 0; j < 8 && full; ++j) {
                uint32_t m = 0;
                const bool f = leafChildren ? leaf_is_full(dag_leaf_ptr(dag, out[j]), dag.M, m)
                                            : header_full(dag_node_ptr(dag, out[j])[0], dag.M, m);
                if (j == 0)
                    fm = m;
                full = f && m == fm;
            }
        }
        rec[0] = full ? header_make_full(0xFFu, fm, dag.M) : header_make(newMask, 0u, dag.M);
        for (uint32_t j = 0; j < 8; ++j)
            rec[1 + j] = j < n ? out[j] : 0u;
        rec[9] = newChild; // outside the item's words: equal items agree on it
        if (full)
            set_flag(kMulti ? &ws.edFull[e * kMaxLevels + L] : &ws.lvl[L].fullCount);
    }
}

// Per level: collapse equal SVO items onto one representative through a
// transient open-addressing table of (hash, index) keys -- contents are
// compared in place, so no slot ever waits on another writer -- and look
// each representative up in the DAG (insertDuplicates / findUnique1 /
// findUnique2, my_gpu_hash_dag_edits.cu:164-246).
template <int G>
__global__ void __launch_bounds__(256) merge_dedup_find(WsDev ws, DagDev dag, uint32_t L)
{
    if (merge_skipped(ws, dag))
        return;
    const LevelDesc me = ws.lvl[L];
    const bool isLeaf = L == dag.leafLevel;
    const uint32_t capMask = dedup_cap(me.count, ws.dedupCap) - 1u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const uint32_t groupsPerBlock = blockDim.x / G;
    for (uint32_t i = blockIdx.x * groupsPerBlock + threadIdx.x / G; i < me.count; i += gridDim.x * groupsPerBlock) {
        const uint32_t w = me.start + i;
        if (isLeaf && ws.created[w] == 2u)
            continue; // unchanged leaf: rep and handle were set by svo_leaf_level
        const uint32_t* rec = isLeaf ? ws.leaves + (size_t)i * ws.leafStride : ws.nodes + (size_t)w * kNodeStride;
        const uint32_t w0 = rec[0], w1 = rec[1];
        const bool empty = isLeaf ? (w0 == 0u && w1 == 0u) : ((w0 & 0xFFu) == 0u);
        if (empty) {
            if (gl == 0)
                ws.rep[w] = kSentinel;
            continue;
        }
        const uint32_t s = isLeaf ? leaf_size_from_words(w0, w1, dag.M) : node_size_from_header(w0);
        Needle<G> n;
        needle_load(n, rec, s, gl);
        const uint32_t h = needle_hash(n, s, gl);
        const unsigned long long key = ((unsigned long long)h << 32) | (unsigned long long)(w + 1u);
        uint32_t p = mix32(h) & capMask;
        uint32_t rep = w;
        while (true) {
            // Equal items share a slot: read first, so duplicates (common
            // surface patterns occur thousands of times) resolve without
            // queueing on one atomic; only an empty slot is claimed by CAS.
            unsigned long long v = 0;
            if (gl == 0) {
                v = ld_filter64(&ws.dedup[p]);
                if (v == 0ull)
                    v = atomicCAS(&ws.dedup[p], 0ull, key);
            }
            const uint32_t vlo = group_bcast<G>((uint32_t)v, 0);
            const uint32_t vhi = group_bcast<G>((uint32_t)(v >> 32), 0);
            if (vlo == 0u && vhi == 0u)
                break;
            if (vhi == h) {
                const uint32_t j = vlo - 1u;
                const uint32_t* other = isLeaf ? ws.leaves + (size_t)(j - me.start) * ws.leafStride
                                               : ws.nodes + (size_t)j * kNodeStride;
                const bool eq = needle_equal_lane(n, other, s, gl);
                if (group_ballot<G>(!eq, gbase) == 0) {
                    rep = j;
                    break;
                }
            }
            p = (p + 1u) & capMask;
        }
        if (gl == 0)
            ws.rep[w] = rep;
        if (rep == w) {
            // A node with a child this merge inserted is new: no find.
            uint32_t r = kNotFound, hint = kNoSlab;
            if (isLeaf || rec[9] == 0u)
                r = table_find_group_hint<G>(dag.tables[s - 2], n, s, h, gl, gbase, hint);
            if (gl == 0) {
                ws.handle[w] = (r == kNotFound) ? kNotFound : (isLeaf ? leaf_handle(dag, r, s) : node_handle(r, s));
                ws.created[w] = 0;
                ws.hint[w] = hint;
            }
        }
    }
}

// Insert the representatives the DAG did not hold (insertUniqueInHashTableAsWarp,
// my_gpu_hash_dag_edits.cu:248-280).  All contents are distinct, so inserts
// never race on equal items, and no find runs in this kernel.
template <int G>
__global__ void __launch_bounds__(256) merge_insert(WsDev ws, DagDev dag, uint32_t L)
{
    if (merge_skipped(ws, dag))
        return;
    __shared__ uint32_t sInserted;
    __shared__ RefArrays sRa;
    if (threadIdx.x == 0)
        sInserted = 0u;
    ref_arrays_stage(sRa, dag);
    __syncthreads();
    const LevelDesc me = ws.lvl[L];
    const bool isLeaf = L == dag.leafLevel;
    const uint8_t role = isLeaf ? 0u : node_role(dag, L);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const uint32_t groupsPerBlock = blockDim.x / G;
    for (uint32_t i = blockIdx.x * groupsPerBlock + threadIdx.x / G; i < me.count; i += gridDim.x * groupsPerBlock) {
        const uint32_t w = me.start + i;
        if (ws.rep[w] != w)
            continue;
        uint32_t hint = ws.hint[w];
        const uint32_t found = ws.handle[w];
        // A found node that lacks the role it is needed in counts as a miss
        // (the reference-count rule in dag_internal.h).  A leaf the edit
        // left unchanged (created 2) is the DAG's own leaf there: plain.
        // The item the DAG holds at this position already has this position's
        // role (unchanged leaves are that case too): no role load.
        if (found != kNotFound) {
            if (found == ws.work[w].w)
                continue;
            if (isLeaf) {
                if (ws.created[w] == 2u
                    || sRa.roles[leaf_handle_size(dag, found) - 2][leaf_handle_ptr(dag, found)] == 0u)
                    continue;
            } else if (sRa.roles[node_handle_size(found) - 2][node_handle_ptr(found)] & role) {
                continue;
            }
        }
        const uint32_t* rec = isLeaf ? ws.leaves + (size_t)i * ws.leafStride : ws.nodes + (size_t)w * kNodeStride;
        const uint32_t w0 = rec[0], w1 = rec[1];
        const uint32_t s = isLeaf ? leaf_size_from_words(w0, w1, dag.M) : node_size_from_header(w0);
        Needle<G> n;
        needle_load(n, rec, s, gl);
        const uint32_t h = needle_hash(n, s, gl);
        if (found != kNotFound) {
            // An equal item stored with the right role may follow in the chain.
            const uint32_t again = table_find_group_hint<G>(dag.tables[s - 2], n, s, h, gl, gbase, hint,
                                                            isLeaf ? kRoleLeafUse : role);
            if (again != kNotFound) {
                if (gl == 0)
                    ws.handle[w] = isLeaf ? leaf_handle(dag, again, s) : node_handle(again, s);
                continue;
            }
        }
        const uint32_t r = table_insert_group<G>(dag.tables[s - 2], n, s, h, gl, gbase, hint);
        if (r != kNotFound)
            refs_on_insert(dag, sRa, n, s, r, role, gl);
        if (gl == 0) {
            if (r == kNotFound) {
                atomicOr(&ws.status->oom, 1u);
                ws.handle[w] = kSentinel;
            } else {
                ws.handle[w] = isLeaf ? leaf_handle(dag, r, s) : node_handle(r, s);
                ws.created[w] = 1;
                atomicAdd(&sInserted, 1u);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && sInserted)
        atomicAdd(dag.stored, (unsigned long long)sInserted);
}

// Publish the new root (my_gpu_hash_dag_edits.cu:626) unless the merge was
// skipped (no leaf, hazard kept from :651) or failed.
__global__ void edit_finish(WsDev ws, DagDev dag, uint32_t* root)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    EditStatus* st = ws.status;
    st->leafCount = ws.lvl[dag.leafLevel].count;
    if (st->overflow || (st->leafCount == 0 && !st->build) || st->oom) {
        st->done = st->overflow == 0 && st->oom == 0;
        st->applied = 0;
        st->rootAfter = st->rootBefore;
        if (!st->done)
            *ws.poison = 1u;
        return;
    }
    const uint32_t r = ws.rep[ws.lvl[0].start];
    const uint32_t newRoot = (r == kSentinel) ? kSentinel : ws.handle[r];
    st->done = 1;
    st->applied = 1;
    st->rootAfter = newRoot;
    *root = newRoot;
}

// ---------------------------------------------------------------- queries
// DAGUtils::get_value (dag_utils.h:146-182).
__global__ void query_voxels_kernel(DagDev dag, const uint32_t* root, const uint32_t* xyz, uint32_t n, uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    uint32_t h = *root;
    uint32_t result = kSentinel;
    if (h != kSentinel) {
        for (uint32_t level = 0; level < dag.leafLevel; ++level) {
            const uint32_t* node = dag_node_ptr(dag, h);
            const uint32_t mask = node[0] & 0xFFu;
            const uint32_t sh = dag.levels - (level + 1);
            const uint32_t c = (((x >> sh) & 1u) << 2) | (((y >> sh) & 1u) << 1) | ((z >> sh) & 1u);
            if (!((mask >> c) & 1u)) {
                h = kSentinel;
                break;
            }
            h = node[1 + __builtin_popcount(mask & ((1u << c) - 1u))];
        }
        if (h != kSentinel) {
            const uint32_t v = ((x & 1u) << 2) | ((y & 1u) << 1) | (z & 1u) | ((x & 2u) << 4) | ((y & 2u) << 3)
                | ((z & 2u) << 2);
            uint32_t m;
            if (leaf_material(dag_leaf_ptr(dag, h), v, dag.M, m))
                result = m;
        }
    }
    out[i] = result;
}

// ---------------------------------------------------------------- terrain
// Seeded 2D Perlin fBm in double precision (only +, -, *, / and floor, so the
// host and device evaluate it bit-identically with contraction off).
struct PerlinDev {
    uint32_t perm[512];
};

__host__ __device__ inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
__host__ __device__ inline double lerp_d(double t, double a, double b) { return a + t * (b - a); }
__host__ __device__ inline double grad2(uint32_t hsh, double x, double z)
{
    switch (hsh & 7u) {
    case 0: return x + z;
    case 1: return -x + z;
    case 2: return x - z;
    case 3: return -x - z;
    case 4: return x;
    case 5: return -x;
    case 6: return z;
    default: return -z;
    }
}
__host__ __device__ inline double perlin2(const uint32_t* P, double x, double z)
{
    const double X = floor(x), Z = floor(z);
    const uint32_t xi = (uint32_t)(int64_t)X & 255u, zi = (uint32_t)(int64_t)Z & 255u;
    const double xf = x - X, zf = z - Z;
    const double u = fade(xf), v = fade(zf);
    const uint32_t aa = P[P[xi] + zi], ab = P[P[xi] + zi + 1], ba = P[P[xi + 1] + zi], bb = P[P[xi + 1] + zi + 1];
    const double x1 = lerp_d(u, grad2(aa, xf, zf), grad2(ba, xf - 1.0, zf));
    const double x2 = lerp_d(u, grad2(ab, xf, zf - 1.0), grad2(bb, xf - 1.0, zf - 1.0));
    return lerp_d(v, x1, x2);
}

__global__ void terrain_heights_kernel(PerlinDev perlin, svdag_terrain tp, uint32_t side, int32_t* out)
{
    const uint64_t n = (uint64_t)side * side;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % side), z = (uint32_t)(i / side);
        double sum = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
        for (uint32_t o = 0; o < tp.octaves; ++o) {
            sum = sum + amp * perlin2(perlin.perm, ((double)x * tp.frequency) * freq, ((double)z * tp.frequency) * freq);
            norm = norm + amp;
            amp = amp * 0.5;
            freq = freq * 2.0;
        }
        const double hgt = tp.base + tp.amplitude * (sum / norm);
        out[i] = (int32_t)floor(hgt);
    }
}

__global__ void pyramid_level_kernel(const int32_t* srcMin, const int32_t* srcMax, uint32_t srcSide, int32_t* dstMin,
                                     int32_t* dstMax)
{
    const uint32_t side = srcSide / 2;
    const uint64_t n = (uint64_t)side * side;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % side), z = (uint32_t)(i / side);
        const uint64_t a = (uint64_t)(2 * z) * srcSide + 2 * x, b = a + srcSide;
        dstMin[i] = min(min(srcMin[a], srcMin[a + 1]), min(srcMin[b], srcMin[b + 1]));
        dstMax[i] = max(max(srcMax[a], srcMax[a + 1]), max(srcMax[b], srcMax[b + 1]));
    }
}

// Largest min-height of a pyramid level: a node of width 2^k at y = 0 is
// solid iff its column block's minimum height is at least 2^k - 1.
__global__ void level_max_kernel(const int32_t* lvl, uint64_t n, int32_t* out)
{
    int32_t m = INT32_MIN;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, lvl[i]);
    for (int o = 32; o >= 1; o >>= 1)
        m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63u) == 0u)
        atomicMax(out, m);
}

// Dense-grid pyramid, level 0: voxel materials (0xFF = empty) to codes;
// a material the DAG cannot hold sets *bad.
__global__ void dense_base_kernel(const uint8_t* in, uint64_t n, uint32_t numMaterials, uint16_t* out, uint32_t* bad)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = in[i];
        if (v == 0xFFu) {
            out[i] = kDenseEmpty;
        } else {
            if (v >= numMaterials)
                *bad = 1u;
            out[i] = (uint16_t)v;
        }
    }
}

// Level k from level k-1: a cell keeps its children's code when all eight
// agree, else it is mixed.
__global__ void dense_level_kernel(const uint16_t* fine, uint32_t fineSide, uint16_t* coarse)
{
    const uint32_t s = fineSide / 2;
    const uint64_t n = (uint64_t)s * s * s;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t z = (uint32_t)(i % s), y = (uint32_t)((i / s) % s), x = (uint32_t)(i / ((uint64_t)s * s));
        const uint16_t c0 = fine[((uint64_t)(2 * x) * fineSide + 2 * y) * fineSide + 2 * z];
        bool same = true;
        for (uint32_t c = 1; c < 8 && same; ++c) {
            const uint32_t fx = 2 * x + ((c >> 2) & 1u), fy = 2 * y + ((c >> 1) & 1u), fz = 2 * z + (c & 1u);
            same = fine[((uint64_t)fx * fineSide + fy) * fineSide + fz] == c0;
        }
        coarse[i] = same ? c0 : kDenseMixed;
    }
}

// splitmix64-driven Fisher-Yates shuffle of 0..255, duplicated to 512 entries.
void perlin_perm(uint64_t seed, uint32_t* perm)
{
    uint32_t p[256];
    for (uint32_t i = 0; i < 256; ++i)
        p[i] = i;
    uint64_t s = seed;
    for (uint32_t i = 255; i >= 1; --i) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t j = (uint32_t)(z % (i + 1));
        std::swap(p[i], p[j]);
    }
    for (uint32_t i = 0; i < 512; ++i)
        perm[i] = p[i & 255u];
}

} // namespace svdag

// ======================================================================== host
namespace svdag {

static uint32_t bit_width(uint32_t v)
{
    uint32_t n = 0;
    while (v) {
        ++n;
        v >>= 1;
    }
    return n;
}

static void ws_release(WsDev& ws)
{
    for (void* p : { (void*)ws.work, (void*)ws.state, (void*)ws.nodes, (void*)ws.leaves, (void*)ws.rep,
                     (void*)ws.handle, (void*)ws.dedup, (void*)ws.lvl, (void*)ws.status, (void*)ws.editOf,
                     (void*)ws.created, (void*)ws.hint, (void*)ws.poison })
        if (p)
            (void)hipFree(p);
    ws = WsDev {};
}

static int ws_alloc(svdag_ctx* ctx, uint64_t capWork, bool withState)
{
    ++ctx->bufferGeneration; // captured wave graphs hold the old pointers
    WsDev old = ctx->ws;
    const bool hadState = old.state != nullptr;
    ws_release(ctx->ws);
    (void)old;
    WsDev& ws = ctx->ws;
    if (capWork > 0x7FFFFFFFull)
        capWork = 0x7FFFFFFFull;
    ws.capWork = (uint32_t)capWork;
    ws.capLeaf = (uint32_t)capWork;
    uint64_t dc = 64;
    while (dc < 2 * capWork)
        dc <<= 1;
    ws.dedupCap = (uint32_t)std::min<uint64_t>(dc, 1ull << 31);
    ws.leafStride = ctx->dag.maxLeafWords;
    SVDAG_HIP(hipMalloc(&ws.work, (size_t)capWork * sizeof(uint4)));
    if (withState || hadState)
        SVDAG_HIP(hipMalloc(&ws.state, (size_t)capWork * sizeof(CopyState)));
    SVDAG_HIP(hipMalloc(&ws.nodes, (size_t)capWork * kNodeStride * 4));
    SVDAG_HIP(hipMalloc(&ws.leaves, (size_t)capWork * ws.leafStride * 4));
    SVDAG_HIP(hipMalloc(&ws.rep, (size_t)capWork * 4));
    SVDAG_HIP(hipMalloc(&ws.handle, (size_t)capWork * 4));
    SVDAG_HIP(hipMalloc(&ws.editOf, (size_t)capWork * 4));
    SVDAG_HIP(hipMalloc(&ws.created, (size_t)capWork));
    SVDAG_HIP(hipMalloc(&ws.hint, (size_t)capWork * 4));
    SVDAG_HIP(hipMalloc(&ws.dedup, (size_t)ws.dedupCap * 8));
    SVDAG_HIP(hipMalloc(&ws.lvl, kMaxLevels * sizeof(LevelDesc)));
    SVDAG_HIP(hipMalloc(&ws.status, sizeof(EditStatus)));
    SVDAG_HIP(hipMalloc(&ws.poison, 4));
    SVDAG_HIP(hipMemsetAsync(ws.poison, 0, 4, ctx->stream));
    SVDAG_HIP(hipMemsetAsync(ws.dedup, 0, (size_t)ws.dedupCap * 8, ctx->stream));
    return SVDAG_OK;
}

// After a table grew: upload the new descriptors and wait for the upload, so
// that no frame trace launched once the caller's FrameQuiesce scope ends can
// read descriptors that still point at freed lines.
int ctx_refresh_tables(svdag_ctx* ctx)
{
    SVDAG_TRY(ctx->ts.upload(ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->dag.tables = ctx->ts.devTables;
    return SVDAG_OK;
}

// Voxel bounds [lo, hi] of an edit: a 2-voxel margin around the editor's
// box, so every node an editor can change intersects them.
Bounds edit_bounds(const svdag_ctx* ctx, const EditorDev& e)
{
    const int64_t side = (int64_t)1 << ctx->dag.levels;
    Bounds b {};
    for (int a = 0; a < 3; ++a) {
        double lo = 0, hi = (double)(side - 1);
        if (e.kind == EK_SPHERE_ADD || e.kind == EK_SPHERE_REMOVE || e.kind == EK_PAINT) {
            const double c = a == 0 ? e.cx : (a == 1 ? e.cy : e.cz);
            lo = c - e.radius - 2.0;
            hi = c + e.radius + 2.0;
        } else if (e.kind == EK_BOX_ADD || e.kind == EK_BOX_REMOVE || e.kind == EK_COPY) {
            lo = e.bmin[a] - 2.0;
            hi = e.bmax[a] + 2.0;
        } else if (e.kind == EK_DENSE) {
            lo = 0.0;
            hi = (double)((1ull << e.denseLog2) - 1);
        }
        // Clamp in double first: far-away editors must not overflow the cast.
        lo = std::min(std::max(lo, -1.0), (double)side);
        hi = std::min(std::max(hi, -1.0), (double)side);
        b.lo[a] = std::min<int64_t>(side - 1, std::max<int64_t>(0, (int64_t)std::floor(lo)));
        b.hi[a] = std::min<int64_t>(side - 1, (int64_t)std::ceil(hi));
        if (b.hi[a] < b.lo[a])
            b.hi[a] = b.lo[a];
    }
    return b;
}

uint64_t level_bound(const svdag_ctx* ctx, const Bounds& b, uint32_t L)
{
    const uint32_t sh = ctx->dag.levels - L;
    uint64_t n = 1;
    for (int a = 0; a < 3; ++a)
        n *= (uint64_t)((b.hi[a] >> sh) - (b.lo[a] >> sh) + 1);
    return n;
}

template <bool kMulti>
static void launch_merge_level(svdag_ctx* ctx, const WsDev& ws, uint32_t L, uint64_t items, uint32_t kCap)
{
    const DagDev& dag = ctx->dag;
    const hipStream_t st = ctx->stream;
    merge_prepare<kMulti><<<grid_for(2 * items, 256, kCap), 256, 0, st>>>(ws, dag, L);
    const uint32_t G = ctx->groupSize;
    const uint32_t gridG = grid_for(items, 256 / G, kCap);
    switch (G) {
    case 4:
        merge_dedup_find<4><<<gridG, 256, 0, st>>>(ws, dag, L);
        merge_insert<4><<<gridG, 256, 0, st>>>(ws, dag, L);
        break;
    case 8:
        merge_dedup_find<8><<<gridG, 256, 0, st>>>(ws, dag, L);
        merge_insert<8><<<gridG, 256, 0, st>>>(ws, dag, L);
        break;
    case 32:
        merge_dedup_find<32><<<gridG, 256, 0, st>>>(ws, dag, L);
        merge_insert<32><<<gridG, 256, 0, st>>>(ws, dag, L);
        break;
    default:
        merge_dedup_find<16><<<gridG, 256, 0, st>>>(ws, dag, L);
        merge_insert<16><<<gridG, 256, 0, st>>>(ws, dag, L);
        break;
    }
}

// Enqueue one edit: ed is the host copy (for grid sizing), edDev the same
// editor in device memory, ws the workspace view whose status / level table
// slots this edit uses.
int launch_edit(svdag_ctx* ctx, const EditorDev& ed, const EditorDev* edDev, const WsDev& ws)
{
    const DagDev& dag = ctx->dag;
    const hipStream_t st = ctx->stream;
    const Bounds b = edit_bounds(ctx, ed);
    const uint32_t kCap = ctx->maxBlocks;
    edit_begin<<<1, 64, 0, st>>>(ws, ctx->dRoot, ed.kind == EK_COPY ? 1u : 0u,
                                 ed.kind == EK_DENSE || ed.kind == EK_TERRAIN ? 1u : 0u);
    for (uint32_t L = 0; L < dag.leafLevel; ++L) {
        const uint64_t items = std::min<uint64_t>(level_bound(ctx, b, L), ws.capWork);
        svo_inner_level<false><<<grid_for(items, 32, kCap), 256, 0, st>>>(ws, dag, edDev, L);
    }
    {
        const uint64_t items = std::min<uint64_t>(level_bound(ctx, b, dag.leafLevel), ws.capLeaf);
        svo_leaf_level<false><<<grid_for(items, 4 * kLeavesPerWave, kCap), 256, 0, st>>>(ws, dag, edDev);
    }
    for (int L = (int)dag.leafLevel; L >= 0; --L)
        launch_merge_level<false>(ctx, ws, (uint32_t)L, std::min<uint64_t>(level_bound(ctx, b, (uint32_t)L), ws.capWork),
                                  kCap);
    edit_finish<<<1, 64, 0, st>>>(ws, dag, ctx->dRoot);
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

int launch_joint_levels(svdag_ctx* ctx, const WsDev& ws, const EditorDev* eds, uint32_t S, const uint64_t* levelItems,
                        uint32_t kCap, hipEvent_t mergeAfter, int part)
{
    const DagDev& dag = ctx->dag;
    const hipStream_t st = ctx->stream;
    if (part != 2) {
        for (uint32_t L = S; L < dag.leafLevel; ++L)
            svo_inner_level<true><<<grid_for(std::min<uint64_t>(levelItems[L], ws.capWork), 32, kCap), 256, 0, st>>>(
                ws, dag, eds, L);
        svo_leaf_level<true><<<grid_for(std::min<uint64_t>(levelItems[dag.leafLevel], ws.capLeaf),
                                        4 * kLeavesPerWave, kCap),
                               256, 0, st>>>(ws, dag, eds);
    }
    // The build above only reads items the root reaches; the merge below
    // finds and inserts, so it waits for whatever mergeAfter marks.
    if (mergeAfter)
        SVDAG_HIP(hipStreamWaitEvent(st, mergeAfter, 0));
    if (part != 1)
        for (int L = (int)dag.leafLevel; L >= (int)S; --L)
            launch_merge_level<true>(ctx, ws, (uint32_t)L, std::min<uint64_t>(levelItems[L], ws.capWork), kCap);
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

// Per-edit device slots for a batch of up to n edits.
static int ensure_batch_slots(svdag_ctx* ctx, uint32_t n)
{
    if (n <= ctx->batchCap)
        return SVDAG_OK;
    // At least 256 edits (a few hundred KB in all), so that batches and
    // edit streams of the usual sizes never reallocate between calls.
    n = std::max<uint32_t>(n, 256u);
    ++ctx->bufferGeneration;
    for (void* p : { (void*)ctx->dEditors, (void*)ctx->dStatusArr, (void*)ctx->dLevelArr, (void*)ctx->dEdCount,
                     (void*)ctx->dEdFull, (void*)ctx->dSeeds, (void*)ctx->dSeedWork, (void*)ctx->dWaves,
                     (void*)ctx->dWaveIdx, ctx->dSpine })
        if (p)
            (void)hipFree(p);
    ctx->dSpine = nullptr;
    for (void* p : { ctx->hBatch, ctx->hFused })
        if (p)
            (void)hipHostFree(p);
    ctx->dEditors = nullptr;
    ctx->hBatch = ctx->hFused = nullptr;
    ctx->batchCap = 0;
    SVDAG_HIP(hipMalloc(&ctx->dEdCount, (size_t)n * kMaxLevels * 4));
    SVDAG_HIP(hipMalloc(&ctx->dEdFull, (size_t)n * kMaxLevels * 4));
    SVDAG_HIP(hipMalloc(&ctx->dSeeds, (size_t)n * sizeof(SeedDesc)));
    SVDAG_HIP(hipMalloc(&ctx->dSeedWork, (size_t)n * 8 * 4));
    SVDAG_HIP(hipMalloc(&ctx->dWaves, (size_t)n * sizeof(uint2)));
    SVDAG_HIP(hipMalloc(&ctx->dWaveIdx, 8)); // wave index, then the parallel spine's done flag
    SVDAG_HIP(hipMalloc(&ctx->dSpine, spine_scratch_bytes()));
    SVDAG_HIP(hipHostMalloc(&ctx->hFused, (size_t)n * (sizeof(SeedDesc) + sizeof(uint2))));
    SVDAG_HIP(hipMalloc(&ctx->dEditors, (size_t)n * sizeof(EditorDev)));
    SVDAG_HIP(hipMalloc(&ctx->dStatusArr, (size_t)n * sizeof(EditStatus)));
    SVDAG_HIP(hipMalloc(&ctx->dLevelArr, (size_t)n * kMaxLevels * sizeof(LevelDesc)));
    SVDAG_HIP(hipHostMalloc(&ctx->hBatch, (size_t)n * (sizeof(EditorDev) + sizeof(EditStatus) + kMaxLevels * sizeof(LevelDesc))));
    ctx->batchCap = n;
    return SVDAG_OK;
}

// Run edits [0, n) in order with no host round trip between them.  In batch
// mode 1 runs of geometric edits go through fused waves (fused.hip); other
// edits (and all edits in mode 0) are enqueued one after another, each
// reading the root the previous one wrote on the device.  On overflow or OOM
// at edit k the root is put back to edit k's input, the workspace or arena
// grows, and edits k.. run again (later results are unreachable).  An edit a
// fused wave hands back runs on its own.
int run_edits(svdag_ctx* ctx, const EditorDev* eds, uint32_t n, svdag_edit_stats* stats, float* totalMs)
{
    SVDAG_TRY(ensure_batch_slots(ctx, n));
    EditorDev* hEd = reinterpret_cast<EditorDev*>(ctx->hBatch);
    EditStatus* hSt = reinterpret_cast<EditStatus*>(hEd + ctx->batchCap);
    LevelDesc* hLv = reinterpret_cast<LevelDesc*>(hSt + ctx->batchCap);
    float ms = 0.0f;
    uint32_t first = 0;
    std::vector<uint8_t> alone(n, 0);
    std::vector<BatchStep> steps;
    std::string why = "none";
    for (int attempt = 0; attempt < 16 && first < n; ++attempt) {
        bool needState = false;
        for (uint32_t k = first; k < n; ++k) {
            hEd[k] = eds[k];
            needState |= eds[k].kind == EK_COPY;
        }
        if (needState && !ctx->ws.state)
            SVDAG_TRY(ws_alloc(ctx, ctx->ws.capWork, true));
        // In the fused modes a single edit also runs as a wave: its levels
        // above the split run in the one-wave spine kernel instead of four
        // launches per level.
        uint32_t S = 0;
        if (ctx->batchMode >= SVDAG_BATCH_FUSED) {
            S = plan_batch(ctx, eds, first, n, steps, alone);
        } else {
            steps.clear();
            for (uint32_t k = first; k < n; ++k)
                steps.push_back({ k, k + 1, false });
        }
        SVDAG_HIP(hipMemcpyAsync(ctx->dEditors + first, hEd + first, (n - first) * sizeof(EditorDev), hipMemcpyHostToDevice,
                                 ctx->stream));
        if (S) {
            const SeedDesc* seeds = reinterpret_cast<const SeedDesc*>(ctx->hFused);
            const uint2* waves = reinterpret_cast<const uint2*>(seeds + ctx->batchCap);
            uint32_t nWaves = 0;
            for (const BatchStep& b : steps)
                nWaves += b.fused ? 1u : 0u;
            SVDAG_HIP(hipMemcpyAsync(ctx->dSeeds + first, seeds + first, (n - first) * sizeof(SeedDesc),
                                     hipMemcpyHostToDevice, ctx->stream));
            SVDAG_HIP(hipMemcpyAsync(ctx->dWaves, waves, nWaves * sizeof(uint2), hipMemcpyHostToDevice, ctx->stream));
            SVDAG_HIP(hipMemsetAsync(ctx->dWaveIdx, 0, 8, ctx->stream));
        }
        SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        for (const BatchStep& b : steps) {
            if (b.fused) {
                SVDAG_TRY(launch_wave(ctx, hEd, b, S));
                continue;
            }
            WsDev ws = ctx->ws;
            ws.status = ctx->dStatusArr + b.first;
            ws.lvl = ctx->dLevelArr + (size_t)b.first * kMaxLevels;
            SVDAG_TRY(launch_edit(ctx, eds[b.first], ctx->dEditors + b.first, ws));
        }
        SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        SVDAG_HIP(hipMemcpyAsync(hSt + first, ctx->dStatusArr + first, (n - first) * sizeof(EditStatus),
                                 hipMemcpyDeviceToHost, ctx->stream));
        SVDAG_HIP(hipMemcpyAsync(hLv + (size_t)first * kMaxLevels, ctx->dLevelArr + (size_t)first * kMaxLevels,
                                 (size_t)(n - first) * kMaxLevels * sizeof(LevelDesc), hipMemcpyDeviceToHost, ctx->stream));
        SVDAG_HIP(hipStreamSynchronize(ctx->stream));
        float part = 0.0f;
        SVDAG_HIP(hipEventElapsedTime(&part, ctx->ev0, ctx->ev1));
        ms += part;
        uint32_t bad = n;
        for (uint32_t k = first; k < n; ++k)
            if (!hSt[k].done) {
                bad = k;
                break;
            }
        for (uint32_t k = first; k < bad; ++k) {
            ctx->seqSpineEdits += hSt[k].seqSpine ? 1u : 0u;
            root_log(ctx);
            ctx->root = hSt[k].rootAfter;
            if (stats) {
                svdag_edit_stats& o = stats[k];
                std::memset(&o, 0, sizeof(o));
                for (uint32_t L = 0; L <= ctx->dag.leafLevel && L < 32; ++L)
                    o.svo_items[L] = hLv[(size_t)k * kMaxLevels + L].count;
                o.applied = hSt[k].applied;
                o.root_before = hSt[k].rootBefore;
                o.root_after = hSt[k].rootAfter;
                o.gpu_ms = part / (float)(n - first);
            }
        }
        if (bad == n) {
            first = n;
            break;
        }
        const EditStatus s = hSt[bad];
        // The failed edit and those after it may have inserted items no root
        // reaches; the next collection marks and recounts.
        ctx->refsDirty = true;
        root_log(ctx);
        ctx->root = s.rootBefore;
        SVDAG_HIP(hipMemcpyAsync(ctx->dRoot, &ctx->root, 4, hipMemcpyHostToDevice, ctx->stream));
        if (s.fallback) {
            ++ctx->handedBackEdits;
            alone[bad] = 1;
            why = "fused spine fallback";
        } else if (s.overflow) {
            uint64_t need = 0;
            for (uint32_t L = 0; L <= ctx->dag.leafLevel; ++L) {
                const LevelDesc& d = hLv[(size_t)bad * kMaxLevels + L];
                need = std::max<uint64_t>(need, (uint64_t)d.start + d.count);
            }
            why = "work items: need " + std::to_string(need) + ", capacity " + std::to_string(ctx->ws.capWork);
            SVDAG_TRY(ws_alloc(ctx, std::max<uint64_t>(2ull * ctx->ws.capWork, need + need / 4), ctx->ws.state != nullptr));
        } else if (s.oom) {
            FrameQuiesce q(ctx);
            for (auto& t : ctx->ts.tables) {
                AllocState a {};
                SVDAG_TRY(table_read_alloc(t, a, ctx->stream));
                if (a.oom) {
                    why = "table of " + std::to_string(t.itemWords) + "-word items full at " + std::to_string(t.slabCap)
                        + " lines (limit " + std::to_string(ctx->slabLimit[t.itemWords - 2]) + ")";
                    SVDAG_TRY(table_grow(
                        t, (uint32_t)std::min<uint64_t>(2ull * t.slabCap, ctx->slabLimit[t.itemWords - 2]), ctx->stream));
                }
            }
            SVDAG_TRY(ctx_refresh_tables(ctx));
        } else {
            set_error("edit " + std::to_string(bad) + " did not complete");
            return SVDAG_E_DEVICE;
        }
        first = bad;
    }
    if (first < n) {
        set_error("edit did not fit after repeated workspace / arena growth (last: " + why + ")");
        return SVDAG_E_CAPACITY;
    }
    if (totalMs)
        *totalMs = ms;
    return SVDAG_OK;
}

// Run one edit to completion.
static int run_edit(svdag_ctx* ctx, const EditorDev& ed, svdag_edit_stats* stats)
{
    float ms = 0.0f;
    SVDAG_TRY(run_edits(ctx, &ed, 1, stats, &ms));
    if (stats)
        stats->gpu_ms = ms;
    return SVDAG_OK;
}

int make_editor(const svdag_ctx* ctx, const svdag_edit_desc& d, EditorDev& e)
{
    e = EditorDev {};
    e.kind = d.kind;
    e.material = d.material;
    if (d.kind > SVDAG_COPY) {
        set_error("unknown edit kind");
        return SVDAG_E_INVALID;
    }
    if (d.material >= ctx->dag.numMaterials) {
        set_error("material does not fit the material bits");
        return SVDAG_E_INVALID;
    }
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(d.center[a]) || !std::isfinite(d.src[a]) || !std::isfinite(d.dst[a])) {
            set_error("non-finite edit coordinates");
            return SVDAG_E_INVALID;
        }
    if (!std::isfinite(d.radius) || d.radius < 0.0f) {
        set_error("radius must be finite and >= 0");
        return SVDAG_E_INVALID;
    }
    if (d.rotation > 3u || (d.rotation != 0u && d.kind != SVDAG_COPY)) {
        set_error("rotation must be 0..3 quarter turns, and only on a copy");
        return SVDAG_E_INVALID;
    }
    // MyGpuSphereEditorBase (my_gpu_dag_editors.h:406-411).
    e.cx = d.center[0] + 0.5f;
    e.cy = d.center[1] + 0.5f;
    e.cz = d.center[2] + 0.5f;
    e.radius = std::max(d.radius, 0.1f);
    e.r2 = e.radius * e.radius;
    // MyGpuBoxEditorBase (50-55); the copy tool centres it at dst + (center - src) (157-164).
    float c[3] = { d.center[0], d.center[1], d.center[2] };
    if (d.kind == SVDAG_COPY) {
        for (int a = 0; a < 3; ++a) {
            c[a] = d.dst[a] + (d.center[a] - d.src[a]);
            e.d2s[a] = -d.dst[a] + d.src[a];
            e.id2s[a] = (int32_t)e.d2s[a];
        }
    }
    for (int a = 0; a < 3; ++a) {
        e.bmin[a] = c[a] - d.radius;
        e.bmax[a] = d.radius == 0.0f ? c[a] + 1.0f : c[a] + d.radius;
    }
    // Rotated copy: the destination voxels overlapping [bmin, bmax) are
    // floor(bmin) .. ceil(bmax) - 1; pivots p = floor(c + 0.5), q = p + id2s.
    e.rot = d.kind == SVDAG_COPY ? d.rotation : 0u;
    if (e.rot) {
        for (int a = 0; a < 3; ++a) {
            if (!(std::fabs(e.bmin[a]) <= 1e9f && std::fabs(e.bmax[a]) <= 1e9f && std::fabs(e.d2s[a]) <= 1e9f)) {
                set_error("copy coordinates out of range");
                return SVDAG_E_INVALID;
            }
            e.ivmin[a] = (int32_t)std::floor(e.bmin[a]);
            e.ivmax[a] = (int32_t)std::ceil(e.bmax[a]) - 1;
            e.pv[a] = (int32_t)std::floor(c[a] + 0.5f);
            e.qv[a] = e.pv[a] + e.id2s[a];
        }
    }
    return SVDAG_OK;
}

// ---------------------------------------------------------------- allocation-free voxel queries
// Staging for svdag_query_voxels_noalloc: the query points (3 words each) of
// at most kQueryScratchPoints points, then one result word per point.  One
// copy per GPU, alive until the process ends; every call on a GPU shares it.
constexpr uint32_t kQueryScratchPoints = 1u << 20;
static __device__ __align__(256) uint32_t scratch[(size_t)kQueryScratchPoints * 4];

// Device address of each GPU's copy of `scratch` (the array's name is only a
// host-side stand-in), filled once per device.
static std::mutex gScratchMu;
static std::unordered_map<int, uint32_t*> gScratchByDevice;

int query_scratch_attach(svdag_ctx* ctx)
{
    int dev = 0;
    SVDAG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(gScratchMu);
    auto it = gScratchByDevice.find(dev);
    if (it == gScratchByDevice.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("query staging: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return SVDAG_E_DEVICE;
        }
        it = gScratchByDevice.emplace(dev, static_cast<uint32_t*>(p)).first;
    }
    ctx->queryScratch = it->second;
    return SVDAG_OK;
}

} // namespace svdag

using namespace svdag;

extern "C" {

int svdag_create(const svdag_config* cfgIn, svdag_ctx** out)
{
    if (!cfgIn || !out) {
        set_error("svdag_create: null argument");
        return SVDAG_E_INVALID;
    }
    svdag_config cfg = *cfgIn;
    if (cfg.levels < 4 || cfg.levels > 20) {
        set_error("levels must be in [4, 20]");
        return SVDAG_E_INVALID;
    }
    if (cfg.material_bits != 0 && cfg.material_bits != 4 && cfg.material_bits != 8) {
        set_error("material_bits must be 0, 4 or 8");
        return SVDAG_E_INVALID;
    }
    if (cfg.group_size != 0 && cfg.group_size != 4 && cfg.group_size != 8 && cfg.group_size != 16 && cfg.group_size != 32) {
        set_error("group_size must be 0, 4, 8, 16 or 32");
        return SVDAG_E_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return SVDAG_E_NO_DEVICE;
    }
    if (cfg.load_factor == 0)
        cfg.load_factor = 75;
    if (cfg.expected_items == 0)
        cfg.expected_items = 1ull << 20;
    if (cfg.workspace_items == 0)
        cfg.workspace_items = 1ull << 22;
    auto* ctx = new svdag_ctx();
    ctx->cfg = cfg;
    ctx->groupSize = cfg.group_size ? cfg.group_size : 16;
    SVDAG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    SVDAG_HIP(hipEventCreate(&ctx->ev0));
    SVDAG_HIP(hipEventCreate(&ctx->ev1));
    SVDAG_HIP(hipEventCreateWithFlags(&ctx->evWait, hipEventDisableTiming));
    // The edit streams' collection stream, result buffers and events for
    // streams of up to 64 batches, made here: creating a stream, events or
    // pinned memory costs milliseconds, which an edit stream would otherwise
    // pay on its first call.
    SVDAG_HIP(hipStreamCreateWithFlags(&ctx->gcStream, hipStreamNonBlocking));
    SVDAG_HIP(hipHostMalloc(&ctx->hStream, gc_counters_bytes() + 64));
    SVDAG_HIP(hipMalloc(&ctx->dStreamFail, 4));
    for (int k = 0; k < 128; ++k) {
        hipEvent_t e = nullptr;
        SVDAG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->streamEvents.push_back(e);
    }
    SVDAG_HIP(hipHostMalloc((void**)&ctx->hStatus, sizeof(EditStatus) + kMaxLevels * sizeof(LevelDesc)));
    ctx->hLevels = reinterpret_cast<LevelDesc*>(ctx->hStatus + 1);

    DagDev& d = ctx->dag;
    d.levels = cfg.levels;
    d.leafLevel = cfg.levels - 2;
    d.M = cfg.material_bits;
    d.numMaterials = 1u << d.M;
    d.minLeafWords = 2 + (d.M ? 1 : 0);
    d.maxLeafWords = 2 + (64 * d.M + 31) / 32;
    d.leafSizeBits = bit_width(d.maxLeafWords - d.minLeafWords);
    const uint32_t maxWords = std::max(9u, d.maxLeafWords);
    d.numTables = maxWords - 2 + 1;
    ctx->slabLimit.resize(d.numTables);
    for (uint32_t i = 0; i < d.numTables; ++i)
        ctx->slabLimit[i] = table_line_limit(2 + i, d.minLeafWords, d.maxLeafWords, d.leafSizeBits);

    // Buckets per table: the table's share of expected_items fills its head
    // lines to load_factor percent.  Half of the items are spread evenly over
    // the tables; the other half follow the node sizes of a surface voxel DAG
    // (a depth-14 fBm terrain: 2..9 words in thousandths below), where nodes
    // with all eight children are the largest group by far.
    static constexpr uint32_t kNodeShare[8] = { 14, 56, 42, 110, 72, 156, 145, 405 };
    ctx->ts.minWords = 2;
    ctx->ts.maxWords = maxWords;
    ctx->ts.tables.resize(d.numTables);
    for (uint32_t i = 0; i < d.numTables; ++i) {
        const double frac = 0.5 / d.numTables + (i < 8 ? 0.5e-3 * kNodeShare[i] : 0.0);
        const uint64_t share = (uint64_t)(frac * (double)cfg.expected_items);
        const uint64_t perLine = (uint64_t)line_slots(2 + i) * cfg.load_factor;
        const uint64_t want = std::max<uint64_t>(64, (share * 100 + perLine - 1) / perLine);
        const uint32_t buckets = (uint32_t)std::min<uint64_t>(want, ctx->slabLimit[i] / 2);
        uint64_t cap = buckets + (cfg.arena_slabs ? cfg.arena_slabs : buckets / 2 + 4096);
        cap = std::min<uint64_t>(cap, ctx->slabLimit[i]);
        SVDAG_TRY(table_alloc(ctx->ts.tables[i], 2 + i, buckets, (uint32_t)cap, ctx->stream, true));
    }
    SVDAG_TRY(ctx_refresh_tables(ctx));
    SVDAG_HIP(hipMalloc(&ctx->dFullyFilled, (size_t)d.levels * d.numMaterials * 4));
    d.fullyFilled = ctx->dFullyFilled;
    SVDAG_HIP(hipMalloc(&ctx->dStored, sizeof(unsigned long long)));
    SVDAG_HIP(hipMemsetAsync(ctx->dStored, 0, sizeof(unsigned long long), ctx->stream));
    d.stored = ctx->dStored;
    SVDAG_HIP(hipMalloc(&ctx->dRoot, 4));
    SVDAG_HIP(hipMemsetAsync(ctx->dRoot, 0xFF, 4, ctx->stream));
    init_fully_filled<<<1, 64, 0, ctx->stream>>>(d);
    SVDAG_HIP(hipGetLastError());
    SVDAG_TRY(ws_alloc(ctx, cfg.workspace_items, false));
    SVDAG_TRY(query_scratch_attach(ctx));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->root = kSentinel;
    *out = ctx;
    return SVDAG_OK;
}

int svdag_destroy(svdag_ctx* ctx)
{
    if (!ctx)
        return SVDAG_OK;
    if (ctx->server)
        (void)svdag_server_stop(ctx, nullptr);
    (void)hipStreamSynchronize(ctx->stream);
    release_wave_graphs(ctx);
    for (uint32_t* p : ctx->gcTouched)
        if (p)
            (void)hipFree(p);
    for (uint8_t* p : ctx->gcMasks)
        if (p)
            (void)hipFree(p);
    if (ctx->dGcMasks)
        (void)hipFree(ctx->dGcMasks);
    ctx->ts.release();
    ws_release(ctx->ws);
    for (void* p : { (void*)ctx->dFullyFilled, (void*)ctx->dRoot, (void*)ctx->hmin, (void*)ctx->hmax,
                     (void*)ctx->gcFront[0], (void*)ctx->gcFront[1], (void*)ctx->gcCounters, ctx->dHits,
                     (void*)ctx->dEditors, (void*)ctx->dStatusArr, (void*)ctx->dLevelArr, (void*)ctx->dEdCount,
                     (void*)ctx->dEdFull, (void*)ctx->dSeeds, (void*)ctx->dSeedWork, (void*)ctx->dWaves,
                     (void*)ctx->dWaveIdx, ctx->dLastFrame, (void*)ctx->dSurf, (void*)ctx->dShade, (void*)ctx->dAo, (void*)ctx->dColor, (void*)ctx->dAccum, ctx->dSpine,
                     (void*)ctx->dStored, (void*)ctx->dGcCand, (void*)ctx->dGcTouchedList })
        if (p)
            (void)hipFree(p);
    for (void* p : { ctx->hBatch, ctx->hFused, ctx->hGcCtr })
        if (p)
            (void)hipHostFree(p);
    (void)hipHostFree(ctx->hStatus);
    if (ctx->hStream)
        (void)hipHostFree(ctx->hStream);
    if (ctx->dStreamFail)
        (void)hipFree(ctx->dStreamFail);
    if (ctx->evGc)
        (void)hipEventDestroy(ctx->evGc);
    if (ctx->gcStream) {
        (void)hipStreamSynchronize(ctx->gcStream);
        (void)hipStreamDestroy(ctx->gcStream);
    }
    for (hipEvent_t e : ctx->streamEvents)
        (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    if (ctx->evWait)
        (void)hipEventDestroy(ctx->evWait);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SVDAG_OK;
}

// The factory build from a voxel source: the grid's material pyramid drives
// the edit pipeline from an empty root, as the terrain's height pyramid does.
int svdag_build_voxels(svdag_ctx* ctx, const uint8_t* materials, uint32_t sideLog2)
{
    if (!ctx || !materials || sideLog2 > ctx->dag.levels || sideLog2 > 11) {
        set_error("svdag_build_voxels: bad arguments (side_log2 must be <= levels and <= 11)");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    const uint64_t side = 1ull << sideLog2;
    const uint64_t n = side * side * side;
    uint64_t off[kMaxLevels] {};
    uint64_t total = 0;
    for (uint32_t k = 0; k <= sideLog2; ++k) {
        off[k] = total;
        const uint64_t sk = side >> k;
        total += sk * sk * sk;
    }
    uint8_t* dIn = nullptr;
    uint16_t* dPyr = nullptr;
    uint32_t* dBad = nullptr;
    auto release = [&]() {
        if (dIn)
            (void)hipFree(dIn);
        if (dPyr)
            (void)hipFree(dPyr);
        if (dBad)
            (void)hipFree(dBad);
    };
    int rc = SVDAG_OK;
    hipError_t e = hipMalloc(&dIn, n);
    if (e == hipSuccess)
        e = hipMalloc(&dPyr, total * sizeof(uint16_t));
    if (e == hipSuccess)
        e = hipMalloc(&dBad, 4);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dIn, materials, n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(dBad, 0, 4, ctx->stream);
    if (e != hipSuccess) {
        set_error(std::string("svdag_build_voxels: ") + hipGetErrorString(e));
        release();
        return SVDAG_E_DEVICE;
    }
    dense_base_kernel<<<grid_for(n, 256, 1u << 16), 256, 0, ctx->stream>>>(dIn, n, ctx->dag.numMaterials, dPyr, dBad);
    for (uint32_t k = 1; k <= sideLog2; ++k) {
        const uint64_t sk = side >> k;
        dense_level_kernel<<<grid_for(sk * sk * sk, 256, 1u << 16), 256, 0, ctx->stream>>>(
            dPyr + off[k - 1], (uint32_t)(side >> (k - 1)), dPyr + off[k]);
    }
    uint32_t bad = 0;
    e = hipMemcpyAsync(&bad, dBad, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_error(std::string("svdag_build_voxels: ") + hipGetErrorString(e));
        release();
        return SVDAG_E_DEVICE;
    }
    if (bad) {
        set_error("svdag_build_voxels: a voxel holds a material the DAG's material bits cannot encode");
        release();
        return SVDAG_E_INVALID;
    }
    // A grid that fills the whole volume with one material is that
    // material's level-0 fully-filled node.  The edit pipeline would build
    // the root from eight filled children with the discovery gate closed
    // (nothing below the root was built), so it is set directly.
    uint16_t top = kDenseMixed;
    e = hipMemcpy(&top, dPyr + off[sideLog2], 2, hipMemcpyDeviceToHost);
    if (e == hipSuccess && sideLog2 == ctx->dag.levels && top < kDenseEmpty) {
        uint32_t root = kSentinel;
        e = hipMemcpy(&root, ctx->dFullyFilled + top, 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(ctx->dRoot, &root, 4, hipMemcpyHostToDevice);
        release();
        if (e != hipSuccess) {
            set_error(std::string("svdag_build_voxels: ") + hipGetErrorString(e));
            return SVDAG_E_DEVICE;
        }
        root_log(ctx);
        ctx->root = root;
        ctx->undo.clear();
        ctx->redo.clear();
        return SVDAG_OK;
    }
    if (e != hipSuccess) {
        set_error(std::string("svdag_build_voxels: ") + hipGetErrorString(e));
        release();
        return SVDAG_E_DEVICE;
    }
    EditorDev ed {};
    ed.kind = EK_DENSE;
    ed.levels = ctx->dag.levels;
    ed.dense = dPyr;
    ed.denseLog2 = sideLog2;
    for (uint32_t k = 0; k <= sideLog2; ++k)
        ed.pyrOffset[k] = off[k];
    // The grid replaces whatever the DAG held: build from the empty root.
    root_log(ctx);
    ctx->root = kSentinel;
    e = hipMemsetAsync(ctx->dRoot, 0xFF, 4, ctx->stream);
    if (e != hipSuccess) {
        set_error(std::string("svdag_build_voxels: ") + hipGetErrorString(e));
        release();
        return SVDAG_E_DEVICE;
    }
    svdag_edit_stats stats;
    rc = run_edit(ctx, ed, &stats);
    (void)hipStreamSynchronize(ctx->stream);
    release();
    ctx->undo.clear();
    ctx->redo.clear();
    return rc;
}

// Terrain through the edit pipeline: a min/max height pyramid classifies the
// nodes (editors.h, EK_TERRAIN).  Below 2^tileLog2 < 2^D the build runs as one
// edit per column tile of that width, so that one edit's work items fit the
// workspace at depth 17; the canonical DAG is the same as long as no
// fully-filled node is wider than a tile (the merge only flags nodes whose
// children sit in the edit that rebuilt them).
static int build_terrain(svdag_ctx* ctx, const svdag_terrain* tp, bool banded, uint32_t bandLog2, uint32_t bandSeed,
                         uint32_t tileLog2)
{
    if (!ctx || !tp || tp->octaves == 0 || tp->octaves > 16 || !std::isfinite(tp->base) || !std::isfinite(tp->amplitude)
        || !std::isfinite(tp->frequency)) {
        set_error("svdag_build_terrain: bad arguments");
        return SVDAG_E_INVALID;
    }
    if (tp->material >= ctx->dag.numMaterials) {
        set_error("terrain material does not fit the material bits");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    const uint32_t D = ctx->dag.levels;
    if (banded && (bandLog2 < 2 || bandLog2 > D)) {
        set_error("terrain band_log2 must lie in [2, levels]");
        return SVDAG_E_INVALID;
    }
    if (tileLog2 == 0)
        tileLog2 = D > 15 ? D - 2 : D;
    if (tileLog2 < 2 || tileLog2 > D) {
        set_error("terrain tile_log2 must lie in [2, levels]");
        return SVDAG_E_INVALID;
    }
    const uint32_t side = 1u << D;
    // Pyramid: level k holds (side >> k)^2 cells.
    uint64_t off = 0;
    for (uint32_t k = 0; k <= D; ++k) {
        ctx->pyrOffset[k] = off;
        off += (uint64_t)(side >> k) * (side >> k);
    }
    for (uint32_t k = D + 1; k < kMaxLevels; ++k)
        ctx->pyrOffset[k] = off;
    if (ctx->hmin)
        (void)hipFree(ctx->hmin);
    if (ctx->hmax)
        (void)hipFree(ctx->hmax);
    ctx->hmin = ctx->hmax = nullptr;
    // Maxima start at level 1 (level 0's are the heights themselves): at depth
    // 17 that saves the 64 GiB a second copy of the heights would take.
    const uint64_t base1 = ctx->pyrOffset[1];
    SVDAG_HIP(hipMalloc(&ctx->hmin, off * 4));
    SVDAG_HIP(hipMalloc(&ctx->hmax, std::max<uint64_t>(off - base1, 1) * 4));
    PerlinDev perlin;
    perlin_perm(tp->seed, perlin.perm);
    terrain_heights_kernel<<<grid_for((uint64_t)side * side, 256, 1u << 16), 256, 0, ctx->stream>>>(perlin, *tp, side, ctx->hmin);
    for (uint32_t k = 1; k <= D; ++k)
        pyramid_level_kernel<<<grid_for((uint64_t)(side >> k) * (side >> k), 256, 1u << 16), 256, 0, ctx->stream>>>(
            ctx->hmin + ctx->pyrOffset[k - 1], k == 1 ? ctx->hmin : ctx->hmax + (ctx->pyrOffset[k - 1] - base1),
            side >> (k - 1), ctx->hmin + ctx->pyrOffset[k], ctx->hmax + (ctx->pyrOffset[k] - base1));
    SVDAG_HIP(hipGetLastError());
    // A tiled build equals the one-edit build only if no fully-filled node is
    // wider than a tile (the merge flags a node only from children rebuilt in
    // the same edit).  The widest solid node sits at y = 0: width 2^k is solid
    // somewhere iff level k's largest minimum height is >= 2^k - 1.  Banded
    // terrains fill nothing wider than a band cube.  Widen the tile to it.
    if (tileLog2 < D) {
        const uint32_t kTop = banded ? std::min(bandLog2, D) : D;
        int32_t* dMax = nullptr;
        SVDAG_HIP(hipMalloc(&dMax, 4));
        for (uint32_t k = kTop; k > tileLog2; --k) {
            const int32_t lo = INT32_MIN;
            int32_t best = INT32_MIN;
            const uint64_t cells = (uint64_t)(side >> k) * (side >> k);
            hipError_t he = hipMemcpyAsync(dMax, &lo, 4, hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess) {
                level_max_kernel<<<grid_for(cells, 256, 4096), 256, 0, ctx->stream>>>(ctx->hmin + ctx->pyrOffset[k],
                                                                                       cells, dMax);
                he = hipGetLastError();
            }
            if (he == hipSuccess)
                he = hipMemcpyAsync(&best, dMax, 4, hipMemcpyDeviceToHost, ctx->stream);
            if (he == hipSuccess)
                he = hipStreamSynchronize(ctx->stream);
            if (he != hipSuccess) {
                (void)hipFree(dMax);
                set_error(std::string("terrain tile check: ") + hipGetErrorString(he));
                return SVDAG_E_DEVICE;
            }
            if ((int64_t)best >= ((int64_t)1 << k) - 1) {
                tileLog2 = k;
                break;
            }
        }
        (void)hipFree(dMax);
    }
    EditorDev e {};
    e.kind = EK_TERRAIN;
    e.material = tp->material;
    e.levels = D;
    e.hmin = ctx->hmin;
    e.hmax = ctx->hmax;
    for (uint32_t k = 0; k < kMaxLevels; ++k)
        e.pyrOffset[k] = ctx->pyrOffset[k];
    e.banded = banded ? 1u : 0u;
    e.bandLog2 = bandLog2;
    e.bandSeed = bandSeed;
    e.materialMask = ctx->dag.numMaterials - 1u;
    // The terrain replaces whatever the DAG held: build from the empty root.
    root_log(ctx);
    ctx->root = kSentinel;
    SVDAG_HIP(hipMemsetAsync(ctx->dRoot, 0xFF, 4, ctx->stream));
    svdag_edit_stats stats;
    if (tileLog2 >= D) {
        SVDAG_TRY(run_edit(ctx, e, &stats));
    } else {
        const uint32_t tiles = 1u << (D - tileLog2), w = 1u << tileLog2;
        e.tiled = 1;
        for (uint32_t tz = 0; tz < tiles; ++tz)
            for (uint32_t tx = 0; tx < tiles; ++tx) {
                e.tileMin[0] = tx * w;
                e.tileMax[0] = tx * w + w;
                e.tileMin[1] = tz * w;
                e.tileMax[1] = tz * w + w;
                SVDAG_TRY(run_edit(ctx, e, &stats));
            }
    }
    ctx->undo.clear();
    ctx->redo.clear();
    return SVDAG_OK;
}

int svdag_build_terrain(svdag_ctx* ctx, const svdag_terrain* tp) { return build_terrain(ctx, tp, false, 0, 0, 0); }

int svdag_build_terrain_banded(svdag_ctx* ctx, const svdag_terrain* tp, uint32_t band_log2, uint32_t band_seed,
                               uint32_t tile_log2)
{
    return build_terrain(ctx, tp, true, band_log2, band_seed, tile_log2);
}

int svdag_terrain_heights(svdag_ctx* ctx, uint32_t x0, uint32_t z0, uint32_t w, uint32_t d, int32_t* out)
{
    if (!ctx || !out || !ctx->hmin) {
        set_error("svdag_terrain_heights: no terrain");
        return SVDAG_E_INVALID;
    }
    const uint32_t side = 1u << ctx->dag.levels;
    if ((uint64_t)x0 + w > side || (uint64_t)z0 + d > side)
        return SVDAG_E_INVALID;
    for (uint32_t z = 0; z < d; ++z)
        SVDAG_HIP(hipMemcpyAsync(out + (size_t)z * w, ctx->hmin + (size_t)(z0 + z) * side + x0, (size_t)w * 4,
                                 hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    return SVDAG_OK;
}

int svdag_root(svdag_ctx* ctx, uint32_t* out)
{
    if (!ctx || !out)
        return SVDAG_E_INVALID;
    *out = ctx->root;
    return SVDAG_OK;
}

static int edit_locked(svdag_ctx* ctx, const svdag_edit_desc* desc, svdag_edit_stats* stats)
{
    EditorDev e;
    SVDAG_TRY(make_editor(ctx, *desc, e));
    // MyGPUHashDAGUndoRedo::add_frame (my_gpu_base_dag.h:23-34), recorded
    // before every edit (my_gpu_hash_dag_edits.cu:638-640).
    if (ctx->undo.empty() || ctx->undo.back() != ctx->root) {
        ctx->redo.clear();
        ctx->undo.push_back(ctx->root);
    }
    svdag_edit_stats local;
    return run_edit(ctx, e, stats ? stats : &local);
}

int svdag_edit(svdag_ctx* ctx, const svdag_edit_desc* desc, svdag_edit_stats* stats)
{
    if (!ctx || !desc)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    return edit_locked(ctx, desc, stats);
}

int svdag_edit_on_stream(svdag_ctx* ctx, const svdag_edit_desc* desc, svdag_edit_stats* stats, void* stream)
{
    if (!ctx || !desc)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (stream) {
        // The edit starts after the work queued on the caller's stream so far.
        SVDAG_HIP(hipEventRecord(ctx->evWait, static_cast<hipStream_t>(stream)));
        SVDAG_HIP(hipStreamWaitEvent(ctx->stream, ctx->evWait, 0));
    }
    return edit_locked(ctx, desc, stats);
}

// Undo frames in edit order, as add_frame before every editMyHashDag.
static void push_undo_frames(svdag_ctx* ctx, const svdag_edit_stats* st, uint32_t n)
{
    for (uint32_t k = 0; k < n; ++k) {
        if (ctx->undo.empty() || ctx->undo.back() != st[k].root_before) {
            ctx->redo.clear();
            ctx->undo.push_back(st[k].root_before);
        }
    }
}

// Edits [0, n) in batches of `batch`, each batch followed by a counting
// collection when collect, all enqueued back to back with one wait at the
// end: the host plans and enqueues every batch while the device runs the
// first ones.  Each collection takes its released roots from its batch's
// device-side results (rc_release_batch).  An edit that cannot complete on
// the device (workspace overflow, arena exhausted, handed back by its wave)
// sets the workspace's poison word, after which the stream's collections
// release nothing; having seen the results, the host runs the rest from that
// edit batch by batch (run_edits, then a collection), as it does a stream
// that is not all fused waves from the start.
static int run_edit_stream(svdag_ctx* ctx, const EditorDev* eds, uint32_t n, uint32_t batch, bool collect,
                           svdag_edit_stats* stats, float* totalMs)
{
    // A batch past the stream is the whole stream (and keeps the batch
    // arithmetic below within 32 bits).
    batch = std::min(batch, n);
    float ms = 0.0f;
    // Batch by batch from edit k0, waiting for each.
    auto serial = [&](uint32_t k0) -> int {
        for (uint32_t a = k0; a < n;) {
            const uint32_t b = std::min<uint32_t>(n, (a / batch + 1) * batch);
            float part = 0.0f;
            SVDAG_TRY(run_edits(ctx, eds + a, b - a, stats + a, &part));
            ms += part;
            if (collect) {
                ctx->undo.clear();
                ctx->redo.clear();
                svdag_gc_stats gs {};
                SVDAG_TRY(gc_run(ctx, &gs, false, false));
            } else {
                push_undo_frames(ctx, stats + a, b - a);
            }
            a = b;
        }
        return SVDAG_OK;
    };
    const uint32_t nb = (n + batch - 1) / batch;
    std::vector<uint32_t> held;
    bool pipe = ctx->batchMode >= SVDAG_BATCH_FUSED;
    if (collect) {
        SVDAG_TRY(gc_settle(ctx));
        for (uint32_t r : ctx->rootLog)
            if (r != kSentinel)
                held.push_back(r);
        std::sort(held.begin(), held.end());
        held.erase(std::unique(held.begin(), held.end()), held.end());
        // A collection that must mark (counts off, or the stored items past
        // 1.5 times the last marking's live count) runs in the serial form.
        const bool markDue = ctx->liveAtMark != 0 && ctx->liveAfterGc > ctx->liveAtMark + ctx->liveAtMark / 2;
        pipe = pipe && !ctx->refsDirty && !markDue && held.size() + batch <= 256;
    }
    std::vector<std::vector<BatchStep>> steps(nb);
    std::vector<uint32_t> S(nb, 0);
    uint32_t nw = 0;
    if (pipe) {
        SVDAG_TRY(ensure_batch_slots(ctx, n));
        const std::vector<uint8_t> alone(n, 0);
        for (uint32_t j = 0; j < nb && pipe; ++j) {
            const uint32_t a = j * batch, b = std::min<uint32_t>(n, a + batch);
            S[j] = plan_batch(ctx, eds, a, b, steps[j], alone, nw);
            pipe = S[j] != 0;
            for (const BatchStep& st : steps[j]) {
                pipe = pipe && st.fused;
                nw += st.fused ? 1u : 0u;
            }
        }
    }
    if (!pipe) {
        SVDAG_TRY(serial(0));
        if (totalMs)
            *totalMs = ms;
        return SVDAG_OK;
    }
    EditorDev* hEd = reinterpret_cast<EditorDev*>(ctx->hBatch);
    EditStatus* hSt = reinterpret_cast<EditStatus*>(hEd + ctx->batchCap);
    LevelDesc* hLv = reinterpret_cast<LevelDesc*>(hSt + ctx->batchCap);
    std::copy(eds, eds + n, hEd);
    if (collect)
        SVDAG_TRY(gc_stream_prepare(ctx));
    const hipStream_t st = ctx->stream;
    const SeedDesc* seeds = reinterpret_cast<const SeedDesc*>(ctx->hFused);
    const uint2* waves = reinterpret_cast<const uint2*>(seeds + ctx->batchCap);
    SVDAG_HIP(hipMemcpyAsync(ctx->dEditors, hEd, n * sizeof(EditorDev), hipMemcpyHostToDevice, st));
    SVDAG_HIP(hipMemcpyAsync(ctx->dSeeds, seeds, n * sizeof(SeedDesc), hipMemcpyHostToDevice, st));
    SVDAG_HIP(hipMemcpyAsync(ctx->dWaves, waves, nw * sizeof(uint2), hipMemcpyHostToDevice, st));
    SVDAG_HIP(hipMemsetAsync(ctx->dWaveIdx, 0, 8, st));
    SVDAG_HIP(hipMemsetAsync(ctx->ws.poison, 0, 4, st));
    SVDAG_HIP(hipMemsetAsync(ctx->dStreamFail, 0, 4, st));
    // Each batch's collection runs on gcStream beside the next batch's first
    // SVO build, which only reads items the root reaches (a collection frees
    // only items no root reaches); that wave's merge, which finds and
    // inserts, waits for the collection.  (HIP-graph waves are replayed as
    // two graphs, the build and the merge, with the wait between them.)
    // Each batch's two events are its own: recording an event again while
    // its last record is still pending can hold the host until it completes.
    const bool overlap = collect && ctx->batchMode >= SVDAG_BATCH_FUSED;
    if (overlap) {
        if (!ctx->gcStream)
            SVDAG_HIP(hipStreamCreateWithFlags(&ctx->gcStream, hipStreamNonBlocking));
        while (ctx->streamEvents.size() < 2 * (size_t)nb) {
            hipEvent_t e = nullptr;
            SVDAG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ctx->streamEvents.push_back(e);
        }
    }
    auto batchDone = [&](uint32_t j) { return ctx->streamEvents[2 * j]; };
    auto gcDone = [&](uint32_t j) { return ctx->streamEvents[2 * j + 1]; };
    SVDAG_HIP(hipEventRecord(ctx->ev0, st));
    void* lastCtr = nullptr;
    const std::vector<uint32_t> none;
    for (uint32_t j = 0; j < nb; ++j) {
        for (size_t w = 0; w < steps[j].size(); ++w)
            SVDAG_TRY(launch_wave(ctx, hEd, steps[j][w], S[j], overlap && j > 0 && w == 0 ? gcDone(j - 1) : nullptr));
        if (collect) {
            const uint32_t a = j * batch, b = std::min<uint32_t>(n, a + batch);
            hipStream_t gs = st;
            if (overlap) {
                SVDAG_HIP(hipEventRecord(batchDone(j), st));
                SVDAG_HIP(hipStreamWaitEvent(ctx->gcStream, batchDone(j), 0));
                gs = ctx->gcStream;
            }
            SVDAG_TRY(gc_enqueue_stream(ctx, a, b - a, j == 0 ? held : none, ctx->dStreamFail, &lastCtr, gs));
            if (overlap)
                SVDAG_HIP(hipEventRecord(gcDone(j), gs));
        }
    }
    if (overlap)
        SVDAG_HIP(hipStreamWaitEvent(st, gcDone(nb - 1), 0));
    SVDAG_HIP(hipEventRecord(ctx->ev1, st));
    SVDAG_HIP(hipMemcpyAsync(hSt, ctx->dStatusArr, n * sizeof(EditStatus), hipMemcpyDeviceToHost, st));
    SVDAG_HIP(hipMemcpyAsync(hLv, ctx->dLevelArr, (size_t)n * kMaxLevels * sizeof(LevelDesc), hipMemcpyDeviceToHost,
                             st));
    uint8_t* hs = static_cast<uint8_t*>(ctx->hStream);
    if (collect) {
        SVDAG_HIP(hipMemcpyAsync(hs, lastCtr, gc_counters_bytes(), hipMemcpyDeviceToHost, st));
        SVDAG_HIP(hipMemcpyAsync(hs + gc_counters_bytes(), ctx->dStreamFail, 4, hipMemcpyDeviceToHost, st));
    }
    SVDAG_HIP(hipStreamSynchronize(st));
    float part = 0.0f;
    SVDAG_HIP(hipEventElapsedTime(&part, ctx->ev0, ctx->ev1));
    ms += part;
    uint32_t bad = n;
    for (uint32_t k = 0; k < n; ++k)
        if (!hSt[k].done) {
            bad = k;
            break;
        }
    // The collections of batches before jb ran: all of them when every edit
    // completed (the last batch may be short), else those before the batch
    // of the first edit that did not.
    const uint32_t jb = bad == n ? nb : bad / batch;
    for (uint32_t k = 0; k < bad; ++k) {
        ctx->seqSpineEdits += hSt[k].seqSpine ? 1u : 0u;
        root_log(ctx);
        ctx->root = hSt[k].rootAfter;
        svdag_edit_stats& o = stats[k];
        std::memset(&o, 0, sizeof(o));
        for (uint32_t L = 0; L <= ctx->dag.leafLevel && L < 32; ++L)
            o.svo_items[L] = hLv[(size_t)k * kMaxLevels + L].count;
        o.applied = hSt[k].applied;
        o.root_before = hSt[k].rootBefore;
        o.root_after = hSt[k].rootAfter;
        o.gpu_ms = part / (float)n;
        const bool batchEnd = k + 1 == n || (k + 1) % batch == 0;
        if (collect && batchEnd && k / batch < jb) {
            ctx->rootLog.clear();
            root_log(ctx);
        }
    }
    if (collect) {
        if (jb > 0) {
            // What svdag_gc does around a collection.
            ctx->undo.clear();
            ctx->redo.clear();
            ctx->traceW = ctx->traceH = 0;
            ++ctx->gcSerial;
            uint32_t fail = 0;
            std::memcpy(&fail, hs + gc_counters_bytes(), 4);
            if (fail)
                ctx->refsDirty = true;
            else if (jb == nb)
                gc_stream_take(ctx, hs, n - (nb - 1) * batch + (nb == 1 ? (uint32_t)held.size() : 0u));
        }
    } else {
        push_undo_frames(ctx, stats, bad);
    }
    if (bad < n) {
        // Later edits may have inserted items no root reaches; the next
        // collection marks and recounts.
        ctx->refsDirty = true;
        SVDAG_HIP(hipMemcpyAsync(ctx->dRoot, &ctx->root, 4, hipMemcpyHostToDevice, st));
        SVDAG_TRY(serial(bad));
    }
    if (totalMs)
        *totalMs = ms;
    return SVDAG_OK;
}

int svdag_edit_batch(svdag_ctx* ctx, const svdag_edit_desc* descs, uint32_t n, svdag_edit_stats* stats, float* gpu_ms)
{
    if (!ctx || (n && !descs))
        return SVDAG_E_INVALID;
    if (n == 0)
        return SVDAG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::vector<EditorDev> eds(n);
    for (uint32_t k = 0; k < n; ++k)
        SVDAG_TRY(make_editor(ctx, descs[k], eds[k]));
    std::vector<svdag_edit_stats> local(stats ? 0 : n);
    svdag_edit_stats* st = stats ? stats : local.data();
    SVDAG_TRY(run_edits(ctx, eds.data(), n, st, gpu_ms));
    push_undo_frames(ctx, st, n);
    return SVDAG_OK;
}

int svdag_edit_stream(svdag_ctx* ctx, const svdag_edit_desc* descs, uint32_t n, uint32_t batch, uint32_t collect,
                      svdag_edit_stats* stats, float* gpu_ms)
{
    if (!ctx || (n && !descs) || batch == 0)
        return SVDAG_E_INVALID;
    if (n == 0)
        return SVDAG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->server) {
        set_error("svdag_edit_stream: a server is running on this context");
        return SVDAG_E_INVALID;
    }
    std::vector<EditorDev> eds(n);
    for (uint32_t k = 0; k < n; ++k)
        SVDAG_TRY(make_editor(ctx, descs[k], eds[k]));
    std::vector<svdag_edit_stats> local(stats ? 0 : n);
    svdag_edit_stats* st = stats ? stats : local.data();
    // No server runs, so no frame traces beside the stream's collections.
    return run_edit_stream(ctx, eds.data(), n, batch, collect != 0, st, gpu_ms);
}

// Times the sample batch under each value of one knob (three runs each, the
// best kept) and leaves the fastest set.  Each run starts from the same root
// and, after a collection, the same tables: the sample's own items are freed
// again, so every timed run inserts them as a real edit would (a run over
// tables that still hold them would time finds that hit instead).
static int tune_knob(svdag_ctx* ctx, const std::vector<EditorDev>& eds, uint32_t* knob, const uint32_t* cand,
                     uint32_t numCand, float* msOut)
{
    const uint32_t n = (uint32_t)eds.size();
    const uint32_t r0 = ctx->root;
    const uint32_t v0 = *knob;
    std::vector<svdag_edit_stats> st(n);
    auto restore = [&]() -> int {
        root_log(ctx);
        ctx->root = r0;
        SVDAG_HIP(hipMemcpyAsync(ctx->dRoot, &ctx->root, 4, hipMemcpyHostToDevice, ctx->stream));
        SVDAG_HIP(hipStreamSynchronize(ctx->stream));
        svdag_gc_stats gs {};
        FrameQuiesce q(ctx);
        return gc_run(ctx, &gs);
    };
    float warm = 0.0f;
    SVDAG_TRY(run_edits(ctx, eds.data(), n, st.data(), &warm));
    SVDAG_TRY(restore());
    uint32_t best = v0;
    float bestMs = 0.0f;
    for (uint32_t c = 0; c < numCand; ++c) {
        *knob = cand[c];
        float m = 0.0f;
        for (int rep = 0; rep < 3; ++rep) {
            float t = 0.0f;
            const int rc = run_edits(ctx, eds.data(), n, st.data(), &t);
            if (rc != SVDAG_OK) {
                *knob = v0;
                (void)restore();
                return rc;
            }
            SVDAG_TRY(restore());
            m = rep == 0 ? t : std::min(m, t);
        }
        msOut[c] = m;
        if (c == 0 || m < bestMs) {
            bestMs = m;
            best = cand[c];
        }
    }
    *knob = best;
    return SVDAG_OK;
}

static int autotune_prepare(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, std::vector<EditorDev>& eds)
{
    eds.resize(n);
    for (uint32_t k = 0; k < n; ++k)
        SVDAG_TRY(make_editor(ctx, sample[k], eds[k]));
    ctx->undo.clear();
    ctx->redo.clear();
    return SVDAG_OK;
}

int svdag_autotune(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, svdag_tune_result* out)
{
    if (!ctx || !out || n == 0 || !sample) {
        set_error("svdag_autotune: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::vector<EditorDev> eds;
    SVDAG_TRY(autotune_prepare(ctx, sample, n, eds));
    static constexpr uint32_t kGroups[4] = { 4, 8, 16, 32 };
    SVDAG_TRY(tune_knob(ctx, eds, &ctx->groupSize, kGroups, 4, out->ms));
    out->group_size = ctx->groupSize;
    return SVDAG_OK;
}

int svdag_autotune_all(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, svdag_tune_report* out)
{
    if (!ctx || !out || n == 0 || !sample) {
        set_error("svdag_autotune_all: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::vector<EditorDev> eds;
    SVDAG_TRY(autotune_prepare(ctx, sample, n, eds));
    *out = svdag_tune_report {};
    static constexpr uint32_t kGroups[4] = { 4, 8, 16, 32 };
    static constexpr uint32_t kBlocks[5] = { 1024, 2048, 4096, 8192, 16384 };
    SVDAG_TRY(tune_knob(ctx, eds, &ctx->groupSize, kGroups, 4, out->ms_group));
    SVDAG_TRY(tune_knob(ctx, eds, &ctx->maxBlocks, kBlocks, 5, out->ms_blocks));
    out->group_size = ctx->groupSize;
    out->max_blocks = ctx->maxBlocks;
    return SVDAG_OK;
}

// MyGPUHashDAG::setLoadFactor (my_gpu_hash_dag.cu:139-155): new tables sized
// for the DAG's own items, and the reachable DAG copied into them.  The
// reference sizes for a number of items per bucket (TARGET_LOAD_FACTOR = 96
// items in 32-slot slabs); loadFactor here is the percentage of head-line
// slots the items fill, as svdag_config.load_factor.  The copy goes through
// the canonical dump and the bulk find-or-insert, after the fully-filled
// chain, as the reference re-adds every item after MyGPUHashDAG::allocate.
// The old tables are kept until the copy succeeded, so a failure leaves the
// DAG as it was.  Handles change: the undo / redo history is cleared and the
// next collection marks.  The caller holds ctx->mu.
static int set_load_factor(svdag_ctx* ctx, uint32_t loadFactor)
{
    DagDev& d = ctx->dag;
    std::vector<uint32_t> dump, counts;
    SVDAG_TRY(canonical_dump(ctx, dump, counts));
    std::vector<uint64_t> per(d.numTables, 0);
    size_t p = 0;
    for (int L = (int)d.leafLevel; L >= 0; --L) {
        const uint32_t count = dump[p + 1];
        p += 2;
        for (uint32_t i = 0; i < count; ++i) {
            ++per[dump[p] - 2];
            p += 1 + dump[p];
        }
    }
    per[d.maxLeafWords - 2] += d.numMaterials;
    per[9 - 2] += (uint64_t)d.numMaterials * d.leafLevel;
    std::vector<HostTable> other(d.numTables);
    auto freeOther = [&]() {
        for (auto& t : other)
            table_free(t);
    };
    for (uint32_t i = 0; i < d.numTables; ++i) {
        const uint64_t slots = line_slots(2 + i);
        const uint64_t want = std::max<uint64_t>(64, (per[i] * 100 + slots * loadFactor - 1) / (slots * loadFactor));
        const uint32_t buckets = (uint32_t)std::min<uint64_t>(want, ctx->slabLimit[i] / 2);
        const uint64_t cap = std::min<uint64_t>((uint64_t)buckets + (per[i] + slots - 1) / slots + buckets / 2 + 4096,
                                                ctx->slabLimit[i]);
        const int rc = table_alloc(other[i], 2 + i, buckets, (uint32_t)cap, ctx->stream, true);
        if (rc != SVDAG_OK) {
            freeOther();
            return rc;
        }
    }
    // No frame trace can read the tables meanwhile: the entry points refuse
    // while a server (the only source of frame traces) runs.
    std::vector<uint32_t> ff((size_t)d.levels * d.numMaterials);
    unsigned long long stored = 0;
    hipError_t he = hipMemcpyAsync(ff.data(), ctx->dFullyFilled, ff.size() * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(&stored, ctx->dStored, sizeof(stored), hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess)
        he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) {
        freeOther();
        set_error(std::string("set_load_factor: ") + hipGetErrorString(he));
        return SVDAG_E_DEVICE;
    }
    const uint32_t oldRoot = ctx->root;
    const std::vector<uint32_t> oldLog = ctx->rootLog;
    const bool oldDirty = ctx->refsDirty;
    release_wave_graphs(ctx);
    ++ctx->bufferGeneration; // captured wave graphs hold the old tables
    ctx->ts.tables.swap(other); // `other` now holds the old tables
    auto attempt = [&]() -> int {
        SVDAG_TRY(ctx_refresh_tables(ctx));
        SVDAG_HIP(hipMemsetAsync(ctx->dStored, 0, sizeof(unsigned long long), ctx->stream));
        init_fully_filled<<<1, 64, 0, ctx->stream>>>(d);
        SVDAG_HIP(hipGetLastError());
        ctx->root = kSentinel;
        ctx->rootLog.clear();
        return import_dump(ctx, dump.data(), dump.size());
    };
    const int rc = attempt();
    if (rc != SVDAG_OK) {
        ctx->ts.tables.swap(other);
        freeOther();
        ctx->root = oldRoot;
        ctx->rootLog = oldLog;
        ctx->refsDirty = oldDirty;
        SVDAG_TRY(ctx_refresh_tables(ctx));
        SVDAG_HIP(hipMemcpyAsync(ctx->dFullyFilled, ff.data(), ff.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        SVDAG_HIP(hipMemcpyAsync(ctx->dStored, &stored, sizeof(stored), hipMemcpyHostToDevice, ctx->stream));
        SVDAG_HIP(hipMemcpyAsync(ctx->dRoot, &ctx->root, 4, hipMemcpyHostToDevice, ctx->stream));
        SVDAG_HIP(hipStreamSynchronize(ctx->stream));
        return rc;
    }
    freeOther();
    // Per-bucket collection bits follow the new bucket counts.
    for (uint32_t*& t : ctx->gcTouched)
        if (t)
            (void)hipFree(t);
    ctx->gcTouched.clear();
    ctx->traceW = ctx->traceH = 0;
    ++ctx->gcSerial; // handles were reissued
    ctx->cfg.load_factor = loadFactor;
    return SVDAG_OK;
}

int svdag_set_load_factor(svdag_ctx* ctx, uint32_t load_factor)
{
    if (!ctx || load_factor < 5 || load_factor > 100) {
        set_error("svdag_set_load_factor: load_factor must lie in [5, 100]");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->server) {
        set_error("svdag_set_load_factor: a server is running on this context");
        return SVDAG_E_INVALID;
    }
    return set_load_factor(ctx, load_factor);
}

int svdag_autotune_load_factor(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, const uint32_t* candidates,
                               uint32_t num_candidates, uint32_t* chosen, float* ms)
{
    if (!ctx || !sample || n == 0 || !candidates || num_candidates == 0 || num_candidates > 8 || !chosen) {
        set_error("svdag_autotune_load_factor: bad arguments");
        return SVDAG_E_INVALID;
    }
    for (uint32_t c = 0; c < num_candidates; ++c)
        if (candidates[c] < 5 || candidates[c] > 100) {
            set_error("svdag_autotune_load_factor: load factors must lie in [5, 100]");
            return SVDAG_E_INVALID;
        }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->server) {
        set_error("svdag_autotune_load_factor: a server is running on this context");
        return SVDAG_E_INVALID;
    }
    std::vector<EditorDev> eds;
    SVDAG_TRY(autotune_prepare(ctx, sample, n, eds));
    uint32_t best = candidates[0];
    float bestMs = 0.0f;
    for (uint32_t c = 0; c < num_candidates; ++c) {
        SVDAG_TRY(set_load_factor(ctx, candidates[c]));
        // The sample timed as one knob value (a warm-up run, then the best
        // of three, each from the same root after a collection).
        uint32_t unused = 0;
        const uint32_t one = 0;
        float m = 0.0f;
        SVDAG_TRY(tune_knob(ctx, eds, &unused, &one, 1, &m));
        if (ms)
            ms[c] = m;
        if (c == 0 || m < bestMs) {
            bestMs = m;
            best = candidates[c];
        }
    }
    if (best != candidates[num_candidates - 1])
        SVDAG_TRY(set_load_factor(ctx, best));
    *chosen = best;
    return SVDAG_OK;
}

int svdag_set_batch_mode(svdag_ctx* ctx, uint32_t mode)
{
    if (!ctx || mode > SVDAG_BATCH_FUSED_GRAPH) {
        set_error("svdag_set_batch_mode: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->batchMode = mode;
    return SVDAG_OK;
}

static int set_root(svdag_ctx* ctx, uint32_t r)
{
    root_log(ctx);
    ctx->root = r;
    SVDAG_HIP(hipMemcpyAsync(ctx->dRoot, &ctx->root, 4, hipMemcpyHostToDevice, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    return SVDAG_OK;
}

int svdag_add_frame(svdag_ctx* ctx, uint32_t root)
{
    if (!ctx)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const bool held = root == ctx->root || std::find(ctx->undo.begin(), ctx->undo.end(), root) != ctx->undo.end()
                      || std::find(ctx->redo.begin(), ctx->redo.end(), root) != ctx->redo.end();
    if (!held) {
        set_error("svdag_add_frame: not a root this context holds");
        return SVDAG_E_INVALID;
    }
    if (!ctx->undo.empty() && ctx->undo.back() == root)
        return SVDAG_OK;
    ctx->redo.clear();
    ctx->undo.push_back(root);
    return SVDAG_OK;
}

int svdag_undo(svdag_ctx* ctx)
{
    if (!ctx)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->undo.empty())
        return SVDAG_OK;
    const uint32_t r = ctx->undo.back();
    ctx->undo.pop_back();
    ctx->redo.push_back(ctx->root);
    return set_root(ctx, r);
}

int svdag_redo(svdag_ctx* ctx)
{
    if (!ctx)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->redo.empty())
        return SVDAG_OK;
    const uint32_t r = ctx->redo.back();
    ctx->redo.pop_back();
    ctx->undo.push_back(ctx->root);
    return set_root(ctx, r);
}

int svdag_history(svdag_ctx* ctx, uint32_t out[2])
{
    if (!ctx || !out)
        return SVDAG_E_INVALID;
    out[0] = (uint32_t)ctx->undo.size();
    out[1] = (uint32_t)ctx->redo.size();
    return SVDAG_OK;
}

int svdag_query_voxels_noalloc(svdag_ctx* ctx, const uint32_t* xyz, uint32_t n, uint32_t* out)
{
    if (!ctx || (n && (!xyz || !out)))
        return SVDAG_E_INVALID;
    if (n > kQueryScratchPoints) {
        set_error("svdag_query_voxels_noalloc: " + std::to_string(n) + " points exceed the staging limit of "
                  + std::to_string(kQueryScratchPoints));
        return SVDAG_E_INVALID;
    }
    if (n == 0)
        return SVDAG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->queryScratch) {
        set_error("svdag_query_voxels_noalloc: no staging address for this context's device");
        return SVDAG_E_DEVICE;
    }
    uint32_t* dIn = ctx->queryScratch;
    uint32_t* dOut = ctx->queryScratch + (size_t)kQueryScratchPoints * 3;
    SVDAG_HIP(hipMemcpyAsync(dIn, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    query_voxels_kernel<<<grid_for(n, 256, 1u << 20), 256, 0, ctx->stream>>>(ctx->dag, ctx->dRoot, dIn, n, dOut);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipMemcpyAsync(out, dOut, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    return SVDAG_OK;
}

int svdag_query_voxels(svdag_ctx* ctx, const uint32_t* xyz, uint32_t n, uint32_t* out)
{
    if (!ctx || (n && (!xyz || !out)))
        return SVDAG_E_INVALID;
    if (n == 0)
        return SVDAG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    uint32_t *dIn = nullptr, *dOut = nullptr;
    SVDAG_HIP(hipMalloc(&dIn, (size_t)n * 12));
    SVDAG_HIP(hipMalloc(&dOut, (size_t)n * 4));
    SVDAG_HIP(hipMemcpyAsync(dIn, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    query_voxels_kernel<<<grid_for(n, 256, 1u << 20), 256, 0, ctx->stream>>>(ctx->dag, ctx->dRoot, dIn, n, dOut);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipMemcpyAsync(out, dOut, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipFree(dIn);
    (void)hipFree(dOut);
    return SVDAG_OK;
}

} // extern "C"
