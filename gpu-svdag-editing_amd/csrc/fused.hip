// Fused multi-edit waves: many pending edits, one SVO build and one merge.
//
// Semantics are those of applying the edits one after another
// (editMyHashDag, my_gpu_hash_dag_edits.cu:633-671, per edit); the fused form
// changes only how the work is scheduled.
//
// Split level S.  Every edit of a batch spans at most two level-S cells per
// axis (S is chosen from the largest edit).  An edit's SVO then has at most
// eight items on every level above S (its "spine"); below S it owns the
// subtrees of its level-S cells.
//
// Waves.  A wave is a run of consecutive edits whose level-S cells are
// disjoint from the voxel bounds of every earlier edit of the wave.  For a
// geometric editor (sphere / box add and remove, sphere paint) a node is
// changed only if it intersects the editor's bounds, and its decision at a
// node depends only on the node and the editor; so the subtrees below S that
// edit k builds and merges are the same whether the earlier edits of the wave
// have run or not.  One wave is therefore
//   fused_seed   for each edit and level-S cell: walk from the wave's input
//                root to the cell and decide it as the parent would
//                (svo_inner_level); subdivided cells become level-S work
//                items tagged with their edit;
//   joint levels the SVO build and merge kernels over levels S .. leaf for all
//                edits of the wave at once, with the fully-filled gate and
//                the no-leaf rule kept per edit;
//   fused_spine  one wave64 replays the edits in order on levels 0 .. S-1:
//                top-down SVO decisions against the current root, then the
//                bottom-up merge (payload cleared, gated full discovery,
//                in-level dedup, DAG find / insert), and publishes each
//                edit's root.
// Copy edits read the live DAG outside their own region and run on their
// own, as does an edit the spine cannot take (more than eight items on a
// level, which the choice of S rules out).
#include "dag_internal.h"
#include <algorithm>
#include <cstring>

namespace svdag {

constexpr uint32_t kSpineItems = 8;
constexpr uint32_t kMaxWave = 256;

// Reset the wave's joint level table and status and its edits' counters.
__global__ void __launch_bounds__(256) wave_begin(WsDev ws, FusedDev f, EditStatus* status)
{
    const uint2 wv = f.waves[*f.waveIdx];
    const uint32_t t = threadIdx.x;
    if (t < kMaxLevels)
        ws.lvl[t] = LevelDesc {};
    if (t == 0)
        *ws.status = EditStatus {};
    for (uint32_t k = wv.x; k < wv.y; ++k) {
        if (t < kMaxLevels) {
            ws.edCount[k * kMaxLevels + t] = 0u;
            ws.edFull[k * kMaxLevels + t] = 0u;
        }
        if (t < 8)
            f.seedWork[k * 8 + t] = kSentinel;
        if (t == 0)
            status[k] = EditStatus {};
    }
}

// One lane per (edit, level-S cell slot).
__global__ void __launch_bounds__(256) fused_seed(WsDev ws, DagDev dag, const EditorDev* eds, FusedDev f,
                                                  const uint32_t* rootPtr)
{
    const uint2 wv = f.waves[*f.waveIdx];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = wv.x + t / 8u, j = t & 7u;
    if (k >= wv.y)
        return;
    const SeedDesc sd = f.seeds[k];
    const uint32_t d[3] = { (j >> 2) & 1u, (j >> 1) & 1u, j & 1u };
    if (d[0] >= sd.n[0] || d[1] >= sd.n[1] || d[2] >= sd.n[2])
        return;
    const uint32_t S = f.S;
    const uint32_t cx = sd.lo[0] + d[0], cy = sd.lo[1] + d[1], cz = sd.lo[2] + d[2];
    uint32_t h = *rootPtr;
    for (uint32_t l = 0; l < S && h != kSentinel; ++l) {
        const uint32_t sh = S - 1u - l;
        const uint32_t c = (((cx >> sh) & 1u) << 2) | (((cy >> sh) & 1u) << 1) | ((cz >> sh) & 1u);
        const uint32_t* node = dag_node_ptr(dag, h);
        const uint32_t mask = node[0] & 0xFFu;
        h = ((mask >> c) & 1u) ? node[1 + __builtin_popcount(mask & ((1u << c) - 1u))] : kSentinel;
    }
    NState st = NS_PARTIAL;
    uint32_t stMat = 0;
    if (h == kSentinel)
        st = NS_EMPTY;
    else if (S != dag.leafLevel && eds[k].kind == EK_PAINT) // only paint reads Filled vs Partial
        st = header_full(dag_node_ptr(dag, h)[0], dag.M, stMat) ? NS_FILLED : NS_PARTIAL;
    uint32_t fillMat = 0;
    if (decide(eds[k], cx, cy, cz, dag.levels - S, st, fillMat) != OP_SUBDIVIDE)
        return;
    const uint32_t g = atomicAdd(&ws.lvl[S].count, 1u);
    atomicAdd(&ws.edCount[k * kMaxLevels + S], 1u);
    if (g < ws.capWork && (S != dag.leafLevel || g < ws.capLeaf)) {
        ws.work[g] = make_uint4(cx, cy, cz, h);
        ws.editOf[g] = k;
        f.seedWork[k * 8 + j] = g;
    } else {
        atomicOr(&ws.status->overflow, 1u);
    }
}

// ------------------------------------------------------------------ parallel spine
// The spine of a wave of sphere and box edits, built one level at a time for
// all edits at once instead of edit after edit.
//
// Why it is the same DAG.  Above S these tools decide by geometry alone (a
// node there is larger than any edit of the wave, so it is neither contained
// nor left untouched by state), so every edit's spine SVO is the same against
// the wave's input root as against the root the edits before it left.  The
// node an edit k builds at spine position P takes each child c either from
// its own level below (c subdivided by k) or from the DAG as the edits
// before k left it: the version of c built by the latest earlier edit whose
// spine contains c, or c's handle in the input DAG.  So level l needs only
// level l + 1 of every edit, and all edits' nodes of one level are built,
// deduplicated and found / inserted together: S sequential steps instead of
// S per edit.  Paint reads node state above S, and anything unexpected (a
// fill above S, more than eight items on a level, a missing seed) leaves the
// wave to the sequential spine.
constexpr uint32_t kParEdits = 64;
constexpr uint32_t kParItems = kParEdits * kSpineItems;
constexpr uint32_t kParThreads = 1024;

struct SpineItem {
    uint32_t x, y, z, orig; // position at its level; its handle in the wave's input DAG
    uint32_t sub;           // children this edit subdivides
    uint32_t dep;           // paint: children whose decision read their state (empty or not)
    uint32_t ref[8];        // per subdivided child: its item one level down, or its seed work item at S
    uint32_t handle, created;
    uint32_t och[8];        // its children in the wave's input DAG (kSentinel: none)
};
struct SpineScratch {
    SpineItem item[kParEdits][kMaxLevels][kSpineItems];
    uint32_t cnt[kParEdits][kMaxLevels];
    uint32_t full[kParEdits][kMaxLevels];
};

size_t spine_scratch_bytes()
{
    return sizeof(SpineScratch);
}

__device__ inline uint32_t joint_handle(const WsDev& ws, uint32_t w, uint32_t& created)
{
    const uint32_t r = ws.rep[w];
    if (r == kSentinel) {
        created = 0u;
        return kSentinel;
    }
    created = ws.created[r] & 1u;
    return ws.handle[r];
}

__global__ void __launch_bounds__(kParThreads) fused_spine_par(WsDev ws, DagDev dag, const EditorDev* eds, FusedDev f,
                                                               uint32_t* rootPtr, EditStatus* status,
                                                               LevelDesc* levels)
{
    __shared__ uint32_t sPos[kParItems][3];
    __shared__ uint32_t sRec[kParItems][kNodeStride];
    __shared__ uint32_t sNew[kParItems], sRep[kParItems], sResH[kParItems], sResC[kParItems];
    __shared__ uint32_t sRepList[kParItems], sSubCnt[kParItems];
    __shared__ uint32_t sSub[kParItems], sChildNew[kParItems], sChildH[kParItems][8];
    __shared__ uint32_t sHash[2 * kParItems], sPosHash[2 * kParItems], sPosOf[kParItems], sPrev[kParItems];
    __shared__ unsigned long long sPosEdits[kParItems];
    __shared__ uint32_t sFullCnt[kParEdits], sDropped[kParEdits];
    __shared__ uint32_t sOk, sBad, sOom, sNumReps;
    __shared__ RefArrays sRa;
    const uint32_t t = threadIdx.x;
    ref_arrays_stage(sRa, dag); // read after the barriers below
    const uint2 wv = f.waves[*f.waveIdx];
    const uint32_t K = wv.y - wv.x, S = f.S;
    const uint32_t leafLevel = dag.leafLevel;
    SpineScratch& sc = *reinterpret_cast<SpineScratch*>(f.spine);
    const EditStatus joint = *ws.status;
    if (t == 0) {
        uint32_t ok = K >= 1 && K <= kParEdits && S >= 1 && S <= leafLevel && !joint.overflow && !joint.oom;
        for (uint32_t k = 0; ok && k < K; ++k) {
            const uint32_t kind = eds[wv.x + k].kind;
            ok = kind == EK_SPHERE_ADD || kind == EK_SPHERE_REMOVE || kind == EK_BOX_ADD || kind == EK_BOX_REMOVE
                || kind == EK_PAINT;
        }
        sOk = ok;
        sBad = 0u;
        sOom = 0u;
    }
    __syncthreads();
    if (!sOk)
        return; // the sequential spine takes the wave
    const uint32_t root0 = *rootPtr;
    // An edit's level may end up without items (an edit off the volume
    // subdivides nothing), so every count starts at zero.
    for (uint32_t q = t; q < K * kMaxLevels; q += blockDim.x)
        sc.cnt[q / kMaxLevels][q % kMaxLevels] = q % kMaxLevels == 0u ? 1u : 0u;
    if (t < K) {
        SpineItem& it = sc.item[t][0][0];
        it.x = it.y = it.z = 0u;
        it.orig = root0;
        sDropped[t] = ws.edCount[(wv.x + t) * kMaxLevels + leafLevel] == 0u ? 1u : 0u;
    }
    __syncthreads();

    // ---- top-down: every edit's spine items on levels 0 .. S-1.
    const uint32_t k = t >> 3, i = t & 7u; // one (edit, item) pair per thread
    for (uint32_t l = 0; l < S; ++l) {
        uint32_t sub = 0, dep = 0, origChild[8];
        const bool active = k < K && i < sc.cnt[k][l];
        if (active) {
            SpineItem& it = sc.item[k][l][i];
            const EditorDev& ed = eds[wv.x + k];
            uint32_t mask = 0;
            const uint32_t* node = nullptr;
            if (it.orig != kSentinel) {
                node = dag_node_ptr(dag, it.orig);
                mask = node[0] & 0xFFu;
            }
            const bool paint = ed.kind == EK_PAINT;
            for (uint32_t c = 0; c < 8; ++c) {
                origChild[c] = ((mask >> c) & 1u) ? node[1 + __builtin_popcount(mask & ((1u << c) - 1u))] : kSentinel;
                const uint32_t cx = (it.x << 1) | ((c >> 2) & 1u), cy = (it.y << 1) | ((c >> 1) & 1u),
                               cz = (it.z << 1) | (c & 1u);
                // Sphere and box decide by geometry alone above S.  Paint
                // reads the child's state (svo_inner_level's NodeDesc), here
                // from the wave's input DAG; the bottom-up checks that the
                // edits before it in the wave left that state as it was.
                NState st = NS_PARTIAL;
                uint32_t stMat = 0;
                if (paint) {
                    if (origChild[c] == kSentinel)
                        st = NS_EMPTY;
                    else if (l + 1u != leafLevel)
                        st = header_full(dag_node_ptr(dag, origChild[c])[0], dag.M, stMat) ? NS_FILLED : NS_PARTIAL;
                }
                uint32_t fm = 0;
                const Op op = decide(ed, cx, cy, cz, dag.levels - (l + 1), st, fm);
                if (op == OP_SUBDIVIDE)
                    sub |= 1u << c;
                else if (!(op == OP_KEEP || (paint && op == OP_EMPTY && st == NS_EMPTY)))
                    atomicOr(&sBad, 1u); // a fill above S: the sequential spine takes the wave
                if (paint && op != OP_KEEP)
                    dep |= 1u << c;
            }
            it.sub = sub;
            it.dep = dep;
            for (uint32_t c = 0; c < 8; ++c)
                it.och[c] = origChild[c];
        }
        if (t < kParItems)
            sSubCnt[t] = active ? (uint32_t)__builtin_popcount(sub) : 0u;
        __syncthreads();
        if (active) {
            SpineItem& it = sc.item[k][l][i];
            uint32_t n = 0;
            for (uint32_t q = 0; q < i; ++q)
                n += sSubCnt[k * kSpineItems + q];
            const SeedDesc sd = f.seeds[wv.x + k];
            for (uint32_t c = 0; c < 8; ++c) {
                if (!((sub >> c) & 1u))
                    continue;
                const uint32_t cx = (it.x << 1) | ((c >> 2) & 1u), cy = (it.y << 1) | ((c >> 1) & 1u),
                               cz = (it.z << 1) | (c & 1u);
                if (l + 1 < S) {
                    if (n < kSpineItems) {
                        SpineItem& ch = sc.item[k][l + 1][n];
                        ch.x = cx;
                        ch.y = cy;
                        ch.z = cz;
                        ch.orig = origChild[c];
                    } else {
                        atomicOr(&sBad, 1u);
                    }
                    it.ref[c] = n++;
                } else {
                    const uint32_t dx = cx - sd.lo[0], dy = cy - sd.lo[1], dz = cz - sd.lo[2];
                    const bool inRange = dx < sd.n[0] && dy < sd.n[1] && dz < sd.n[2];
                    const uint32_t w = inRange ? f.seedWork[(wv.x + k) * 8 + ((dx << 2) | (dy << 1) | dz)] : kSentinel;
                    if (w == kSentinel)
                        atomicOr(&sBad, 1u);
                    it.ref[c] = w;
                }
            }
            if (i == 0 && l + 1 < S) {
                uint32_t tot = 0;
                for (uint32_t q = 0; q < kSpineItems; ++q)
                    tot += sSubCnt[k * kSpineItems + q];
                sc.cnt[k][l + 1] = tot < kSpineItems ? tot : kSpineItems;
                if (tot > kSpineItems)
                    atomicOr(&sBad, 1u);
            }
        }
        __syncthreads();
        if (sBad)
            return; // nothing written outside the scratch: the sequential spine takes the wave
    }

    // ---- bottom-up: levels S-1 .. 0, every edit's nodes of a level at once.
    const uint32_t lane = t & 63u, gl = t & 7u, gbase = lane & ~7u;
    for (int l = (int)S - 1; l >= 0; --l) {
        const bool active = k < K && i < sc.cnt[k][l] && !sDropped[k];
        const uint32_t idx = k * kSpineItems + i;
        const bool childLevelS = (uint32_t)l + 1 == S;
        const bool leafChildren = (uint32_t)l + 1 == leafLevel;
        // Stage the level in LDS: positions, and the subdivided children with
        // their merge results.  A skipped (no-leaf) edit rebuilds nothing, so
        // it stages no subdivided children.
        if (t < kParItems) {
            uint32_t subm = 0, newm = 0;
            if (k < K && i < sc.cnt[k][l]) {
                const SpineItem& it = sc.item[k][l][i];
                sPos[t][0] = it.x;
                sPos[t][1] = it.y;
                sPos[t][2] = it.z;
                if (!sDropped[k]) {
                    subm = it.sub;
                    for (uint32_t c = 0; c < 8; ++c) {
                        if (!((subm >> c) & 1u))
                            continue;
                        uint32_t cr = 0;
                        if (childLevelS) {
                            sChildH[t][c] = joint_handle(ws, it.ref[c], cr);
                        } else {
                            const SpineItem& ch = sc.item[k][l + 1][it.ref[c]];
                            sChildH[t][c] = ch.handle;
                            cr = ch.created;
                        }
                        newm |= cr << c;
                    }
                }
            } else {
                sPos[t][0] = kSentinel;
            }
            sSub[t] = subm;
            sChildNew[t] = newm;
        }
        for (uint32_t h = t; h < 2 * kParItems; h += blockDim.x) {
            sHash[h] = 0u;
            sPosHash[h] = 0u;
        }
        if (t < kParItems)
            sPosEdits[t] = 0ull;
        if (t < kParEdits)
            sFullCnt[t] = 0u;
        if (t == 0)
            sNumReps = 0u;
        __syncthreads();
        // Group the level's items by position: the first item to claim a
        // position names it, and collects the set of edits with an item there.
        if (t < kParItems && sPos[t][0] != kSentinel) {
            const uint32_t px = sPos[t][0], py = sPos[t][1], pz = sPos[t][2];
            uint32_t p = mix32(px * 0x9E3779B1u ^ mix32(py * 0x85EBCA77u ^ mix32(pz))) & (2 * kParItems - 1u);
            uint32_t owner = t;
            while (true) {
                uint32_t v = sPosHash[p];
                if (v == 0u) {
                    v = atomicCAS(&sPosHash[p], 0u, t + 1u);
                    if (v == 0u)
                        break;
                }
                if (sPos[v - 1u][0] == px && sPos[v - 1u][1] == py && sPos[v - 1u][2] == pz) {
                    owner = v - 1u;
                    break;
                }
                p = (p + 1u) & (2 * kParItems - 1u);
            }
            sPosOf[t] = owner;
            atomicOr(&sPosEdits[owner], 1ull << (t / kSpineItems));
        }
        __syncthreads();
        // Each item's predecessor: the same position's item in the latest
        // earlier edit that has one.
        if (t < kParItems) {
            uint32_t prev = kSentinel;
            if (sPos[t][0] != kSentinel) {
                const uint32_t e = t / kSpineItems;
                const unsigned long long before = sPosEdits[sPosOf[t]] & ((1ull << e) - 1ull);
                if (before != 0ull) {
                    const uint32_t j = 63u - (uint32_t)__builtin_clzll(before);
                    for (uint32_t q = 0; q < kSpineItems; ++q) {
                        const uint32_t o = j * kSpineItems + q;
                        if (sPos[o][0] != kSentinel && sPosOf[o] == sPosOf[t]) {
                            prev = o;
                            break;
                        }
                    }
                }
            }
            sPrev[t] = prev;
        }
        __syncthreads();
        bool empty = true;
        if (active) {
            const SpineItem& it = sc.item[k][l][i];
            const uint32_t mySub = sSub[idx];
            uint32_t ver[8];
            for (uint32_t c = 0; c < 8; ++c)
                ver[c] = ((mySub >> c) & 1u) ? sChildH[idx][c] : kSentinel;
            // A child this edit leaves alone is the version the latest earlier
            // edit rebuilt, else the input DAG's: walk the position's
            // predecessors until every such child has its version.
            uint32_t left = ~mySub & 0xFFu, from = 0u;
            for (uint32_t o = sPrev[idx]; left != 0u && o != kSentinel; o = sPrev[o]) {
                const uint32_t hit = sSub[o] & left;
                for (uint32_t c = 0; c < 8; ++c)
                    if ((hit >> c) & 1u)
                        ver[c] = sChildH[o][c];
                left &= ~hit;
                from |= hit;
            }
            from |= mySub;
            uint32_t outH[8], nOut = 0, newMask = 0;
            const uint32_t newChild = sChildNew[idx] != 0u ? 1u : 0u;
            for (uint32_t c = 0; c < 8; ++c) {
                const uint32_t h = ((from >> c) & 1u) ? ver[c] : it.och[c];
                if (h != kSentinel) {
                    outH[nOut++] = h;
                    newMask |= 1u << c;
                }
            }
            // Discovery, gated on this edit's full items one level down.
            const uint32_t prevFull = childLevelS ? ws.edFull[(wv.x + k) * kMaxLevels + S] : sc.full[k][l + 1];
            bool full = false;
            uint32_t fm = 0;
            if (prevFull != 0u && newMask == 0xFFu) {
                full = true;
                for (uint32_t j = 0; j < 8 && full; ++j) {
                    uint32_t m = 0;
                    const bool fj = leafChildren ? leaf_is_full(dag_leaf_ptr(dag, outH[j]), dag.M, m)
                                                 : header_full(dag_node_ptr(dag, outH[j])[0], dag.M, m);
                    if (j == 0)
                        fm = m;
                    full = fj && m == fm;
                }
            }
            empty = newMask == 0u;
            uint32_t* rec = sRec[idx];
            rec[0] = full ? header_make_full(0xFFu, fm, dag.M) : header_make(newMask, 0u, dag.M);
            for (uint32_t j = 0; j < 8; ++j)
                rec[1 + j] = j < nOut ? outH[j] : 0u;
            rec[9] = 0u;
            sNew[idx] = newChild;
            if (full)
                atomicAdd(&sFullCnt[k], 1u);
        }
        // A paint item decided its state-dependent children on the input
        // DAG's state; the edits before it in the wave may have made such a
        // child empty or non-empty.  The DAG would come out the same (the
        // paint's own cells are disjoint from those edits), but its SVO would
        // not, so such a wave goes to the sequential spine, which decides on
        // the current state.  Nothing of the wave is published before the
        // last level, and what was inserted below is what the sequential
        // spine builds again and finds.
        if (k < K && i < sc.cnt[k][l] && eds[wv.x + k].kind == EK_PAINT) {
            const SpineItem& it = sc.item[k][l][i];
            uint32_t need = it.dep, prior[8];
            for (uint32_t c = 0; c < 8; ++c)
                prior[c] = it.och[c];
            for (uint32_t o = sPrev[idx]; need != 0u && o != kSentinel; o = sPrev[o]) {
                const uint32_t hit = sSub[o] & need;
                for (uint32_t c = 0; c < 8; ++c)
                    if ((hit >> c) & 1u)
                        prior[c] = sChildH[o][c];
                need &= ~hit;
            }
            for (uint32_t c = 0; c < 8; ++c)
                if (((it.dep >> c) & 1u) && (prior[c] == kSentinel) != (it.och[c] == kSentinel))
                    atomicOr(&sBad, 1u);
        }
        __syncthreads();
        if (sBad)
            return; // the sequential spine takes the wave
        // In-level dedup over every edit's nodes: the first claimant of a
        // content is its representative.
        if (active && !empty) {
            const uint32_t* rec = sRec[idx];
            uint32_t hh = 0;
            for (uint32_t w = 0; w < kNodeStride; ++w)
                hh = mix32(hh ^ (rec[w] + w * 0x9E3779B9u));
            uint32_t p = hh & (2 * kParItems - 1u);
            uint32_t rep = idx;
            while (true) {
                uint32_t v = sHash[p];
                if (v == 0u) {
                    v = atomicCAS(&sHash[p], 0u, idx + 1u);
                    if (v == 0u)
                        break;
                }
                bool eq = true;
                for (uint32_t w = 0; w < kNodeStride && eq; ++w)
                    eq = sRec[v - 1u][w] == rec[w];
                if (eq) {
                    rep = v - 1u;
                    break;
                }
                p = (p + 1u) & (2 * kParItems - 1u);
            }
            sRep[idx] = rep;
            if (rep == idx)
                sRepList[atomicAdd(&sNumReps, 1u)] = idx;
        } else if (t < kParItems) {
            sRep[t] = kSentinel;
        }
        if (t < kParEdits && t < K)
            sc.full[t][l] = sFullCnt[t];
        __syncthreads();
        // DAG finds of the representatives, eight lanes each, then the
        // inserts of those not found.  Finds and inserts never run at once:
        // an insert claims its slot's tag before writing the item, so a
        // concurrent find could otherwise match a half-written slot.
        const uint32_t nReps = sNumReps;
        for (uint32_t r = t >> 3; r < nReps; r += kParThreads / 8u) {
            const uint32_t ri = sRepList[r];
            const uint32_t* rec = sRec[ri];
            const uint32_t s = node_size_from_header(rec[0]);
            Needle<8> nd;
            needle_load(nd, rec, s, gl);
            const uint32_t hsh = needle_hash(nd, s, gl);
            uint32_t res = kNotFound, hint = kNoSlab;
            if (sNew[ri] == 0u)
                res = table_find_group_hint<8>(dag.tables[s - 2], nd, s, hsh, gl, gbase, hint,
                                               node_role(dag, (uint32_t)l));
            if (gl == 0) {
                sResH[ri] = res == kNotFound ? hint : node_handle(res, s);
                sResC[ri] = res == kNotFound ? 2u : 0u; // 2: to insert, from the line in sResH
            }
        }
        __syncthreads();
        for (uint32_t r = t >> 3; r < nReps; r += kParThreads / 8u) {
            const uint32_t ri = sRepList[r];
            if (sResC[ri] != 2u)
                continue; // uniform over the eight lanes of the item
            const uint32_t* rec = sRec[ri];
            const uint32_t s = node_size_from_header(rec[0]);
            Needle<8> nd;
            needle_load(nd, rec, s, gl);
            const uint32_t hsh = needle_hash(nd, s, gl);
            const uint32_t res = table_insert_group<8>(dag.tables[s - 2], nd, s, hsh, gl, gbase, sResH[ri]);
            if (res != kNotFound) {
                refs_on_insert(dag, sRa, nd, s, res, node_role(dag, (uint32_t)l), gl);
                if (gl == 0)
                    atomicAdd(dag.stored, 1ull);
            }
            // Every lane has read sResH above; the group's lanes are in
            // lockstep within the wave, so lane 0 may overwrite it.
            if (gl == 0) {
                if (res == kNotFound) {
                    sOom = 1u;
                    sResH[ri] = kSentinel;
                } else {
                    sResH[ri] = node_handle(res, s);
                }
                sResC[ri] = 1u;
            }
        }
        __syncthreads();
        if (active) {
            SpineItem& it = sc.item[k][l][i];
            const uint32_t rep = sRep[idx];
            it.handle = rep == kSentinel ? kSentinel : sResH[rep];
            it.created = rep == kSentinel ? 0u : sResC[rep];
        }
        __syncthreads();
        if (sOom)
            break;
    }

    // ---- per-edit results, in edit order.
    if (sOom) {
        if (t == 0) {
            EditStatus st {};
            st.oom = 1u;
            st.rootBefore = st.rootAfter = root0;
            status[wv.x] = st;
            *ws.poison = 1u;
            *f.spineDone = 1u;
        }
        return;
    }
    for (uint32_t q = t; q < K * kMaxLevels; q += blockDim.x) {
        const uint32_t e = q / kMaxLevels, lev = q % kMaxLevels;
        LevelDesc d {};
        d.count = lev < S ? sc.cnt[e][lev] : ws.edCount[(wv.x + e) * kMaxLevels + lev];
        levels[(wv.x + e) * kMaxLevels + lev] = d;
    }
    if (t == 0) {
        uint32_t root = root0;
        for (uint32_t e = 0; e < K; ++e) {
            EditStatus st {};
            st.done = 1u;
            st.rootBefore = root;
            if (!sDropped[e]) {
                st.applied = 1u;
                st.leafCount = ws.edCount[(wv.x + e) * kMaxLevels + leafLevel];
                root = sc.item[e][0][0].handle;
            }
            st.rootAfter = root;
            status[wv.x + e] = st;
        }
        *rootPtr = root;
        *f.spineDone = 1u;
    }
}

// One wave64 (the whole block) replays the wave's edits on levels 0 .. S-1.
__global__ void __launch_bounds__(64) fused_spine(WsDev ws, DagDev dag, const EditorDev* eds, FusedDev f,
                                                  uint32_t* rootPtr, EditStatus* status, LevelDesc* levels)
{
    __shared__ uint4 sItem[kMaxLevels][kSpineItems];
    __shared__ uint32_t sRec[kMaxLevels][kSpineItems][kNodeStride];
    __shared__ uint32_t sHandle[kMaxLevels][kSpineItems];
    __shared__ uint32_t sCreated[kMaxLevels][kSpineItems]; // 1: the spine inserted it
    __shared__ uint32_t sHasNew[kSpineItems];               // item has a created child
    __shared__ uint32_t sCnt[kMaxLevels];
    __shared__ RefArrays sRa;
    const uint32_t lane = threadIdx.x;
    ref_arrays_stage(sRa, dag); // read after the barriers below
    if (*f.spineDone) {
        // The parallel spine took this wave.
        __syncthreads();
        if (lane == 0) {
            *f.spineDone = 0u;
            *f.waveIdx += 1u;
        }
        return;
    }
    const uint2 wv = f.waves[*f.waveIdx];
    const uint32_t S = f.S;
    const uint32_t leafLevel = dag.leafLevel;
    uint32_t root = *rootPtr;
    const EditStatus joint = *ws.status;
    if (joint.overflow || joint.oom) {
        // Nothing of the wave is applied; the host grows what ran out and
        // runs the edits again from the wave's first one.
        if (lane == 0) {
            EditStatus st {};
            st.overflow = joint.overflow;
            st.oom = joint.oom;
            st.rootBefore = st.rootAfter = root;
            status[wv.x] = st;
            *ws.poison = 1u;
        }
        if (lane < kMaxLevels)
            levels[wv.x * kMaxLevels + lane] = ws.lvl[lane];
        __syncthreads();
        if (lane == 0)
            *f.waveIdx += 1u;
        return;
    }
    const uint32_t i8 = lane >> 3, c8 = lane & 7u, gbase = lane & ~7u;
    for (uint32_t k = wv.x; k < wv.y; ++k) {
        const EditorDev& ed = eds[k];
        const SeedDesc sd = f.seeds[k];
        // ---- top-down: the edit's SVO on levels 0 .. S-1 (svo_inner_level).
        if (lane == 0) {
            sItem[0][0] = make_uint4(0u, 0u, 0u, root);
            sCnt[0] = 1u;
        }
        __syncthreads();
        bool bad = false;
        for (uint32_t l = 0; l < S; ++l) {
            const uint32_t n = sCnt[l];
            const bool active = i8 < n;
            const uint32_t childLevel = l + 1u;
            const bool leafChildren = childLevel == leafLevel;
            const uint4 item = active ? sItem[l][i8] : make_uint4(0u, 0u, 0u, kSentinel);
            uint32_t childHandle = kSentinel;
            if (item.w != kSentinel) {
                const uint32_t* node = dag_node_ptr(dag, item.w);
                const uint32_t mask = node[0] & 0xFFu;
                if ((mask >> c8) & 1u)
                    childHandle = node[1 + __builtin_popcount(mask & ((1u << c8) - 1u))];
            }
            const uint32_t cx = (item.x << 1) | ((c8 >> 2) & 1u);
            const uint32_t cy = (item.y << 1) | ((c8 >> 1) & 1u);
            const uint32_t cz = (item.z << 1) | (c8 & 1u);
            NState st = NS_PARTIAL;
            uint32_t stMat = 0;
            if (childHandle == kSentinel)
                st = NS_EMPTY;
            else if (!leafChildren && ed.kind == EK_PAINT) // only paint reads Filled vs Partial
                st = header_full(dag_node_ptr(dag, childHandle)[0], dag.M, stMat) ? NS_FILLED : NS_PARTIAL;
            uint32_t fillMat = 0;
            const Op op = active ? decide(ed, cx, cy, cz, dag.levels - childLevel, st, fillMat) : OP_KEEP;
            const bool sub = active && op == OP_SUBDIVIDE;
            const unsigned long long subBal = __ballot(sub);
            const uint32_t gsub = (uint32_t)(subBal >> gbase) & 0xFFu;
            uint32_t ref = childHandle;
            if (sub) {
                if (childLevel < S) {
                    const uint32_t idx = (uint32_t)__builtin_popcountll(subBal & ((1ull << lane) - 1ull));
                    if (idx < kSpineItems)
                        sItem[childLevel][idx] = make_uint4(cx, cy, cz, childHandle);
                    ref = idx;
                } else {
                    const uint32_t dx = cx - sd.lo[0], dy = cy - sd.lo[1], dz = cz - sd.lo[2];
                    const bool inRange = dx < sd.n[0] && dy < sd.n[1] && dz < sd.n[2];
                    ref = inRange ? f.seedWork[k * 8 + ((dx << 2) | (dy << 1) | dz)] : kSentinel;
                    bad |= ref == kSentinel;
                }
            } else if (op == OP_EMPTY) {
                ref = kSentinel;
            } else if (op == OP_FILL) {
                ref = dag.fullyFilled[childLevel * dag.numMaterials + fillMat];
            }
            const bool out = active && ref != kSentinel;
            const uint32_t cmask = (uint32_t)(__ballot(out) >> gbase) & 0xFFu;
            if (active) {
                uint32_t* rec = sRec[l][i8];
                const uint32_t nOut = __builtin_popcount(cmask);
                if (out)
                    rec[1 + __builtin_popcount(cmask & ((1u << c8) - 1u))] = ref;
                if (c8 >= nOut)
                    rec[1 + c8] = 0u;
                if (c8 == 0) {
                    rec[0] = cmask ? header_make(cmask, gsub, dag.M) : 0u;
                    rec[9] = 0u;
                }
            }
            const uint32_t total = (uint32_t)__builtin_popcountll(subBal);
            if (childLevel < S && total > kSpineItems)
                bad = true;
            if (lane == 0 && childLevel < S)
                sCnt[childLevel] = min(total, kSpineItems);
            __syncthreads();
        }
        if (__ballot(bad) != 0) {
            // Hand the edit back to the host, which runs it on its own.
            if (lane == 0) {
                EditStatus st {};
                st.fallback = 1u;
                st.rootBefore = st.rootAfter = root;
                status[k] = st;
                *ws.poison = 1u;
            }
            break;
        }
        if (lane < kMaxLevels) {
            LevelDesc d {};
            d.count = lane < S ? sCnt[lane] : ws.edCount[k * kMaxLevels + lane];
            levels[k * kMaxLevels + lane] = d;
        }
        const uint32_t leafCount = ws.edCount[k * kMaxLevels + leafLevel];
        if (leafCount == 0) {
            // No leaf: the edit is dropped (my_gpu_hash_dag_edits.cu:651).
            if (lane == 0) {
                EditStatus st {};
                st.done = 1u;
                st.seqSpine = 1u;
                st.rootBefore = st.rootAfter = root;
                status[k] = st;
            }
            __syncthreads();
            continue;
        }
        // ---- bottom-up merge of levels S-1 .. 0.
        uint32_t prevFull = ws.edFull[k * kMaxLevels + S];
        bool oom = false;
        for (int l = (int)S - 1; l >= 0; --l) {
            const uint32_t n = sCnt[l];
            const bool childSeed = (uint32_t)l + 1u == S;
            const bool leafChildren = (uint32_t)l + 1u == leafLevel;
            // updateChildPointers + gated discover_fully_filled: one lane per item.
            bool full = false;
            if (lane < n) {
                uint32_t* rec = sRec[l][lane];
                const uint32_t hdr = rec[0];
                const uint32_t mask = hdr & 0xFFu;
                if (mask) {
                    const uint32_t imask = header_payload(hdr, dag.M) & 0xFFu;
                    uint32_t outH[8];
                    uint32_t nOut = 0, newMask = 0, q = 0, newChild = 0;
                    for (uint32_t c = 0; c < 8; ++c) {
                        if (!((mask >> c) & 1u))
                            continue;
                        uint32_t h = rec[1 + q++];
                        if ((imask >> c) & 1u) {
                            if (childSeed) {
                                const uint32_t r = ws.rep[h];
                                h = (r == kSentinel) ? kSentinel : ws.handle[r];
                                if (r != kSentinel)
                                    newChild |= ws.created[r] & 1u;
                            } else {
                                newChild |= sCreated[l + 1][h];
                                h = sHandle[l + 1][h];
                            }
                        }
                        if (h != kSentinel) {
                            outH[nOut++] = h;
                            newMask |= 1u << c;
                        }
                    }
                    uint32_t fm = 0;
                    if (prevFull != 0 && newMask == 0xFFu) {
                        full = true;
                        for (uint32_t j = 0; j < 8 && full; ++j) {
                            uint32_t m = 0;
                            const bool fj = leafChildren ? leaf_is_full(dag_leaf_ptr(dag, outH[j]), dag.M, m)
                                                         : header_full(dag_node_ptr(dag, outH[j])[0], dag.M, m);
                            if (j == 0)
                                fm = m;
                            full = fj && m == fm;
                        }
                    }
                    rec[0] = full ? header_make_full(0xFFu, fm, dag.M) : header_make(newMask, 0u, dag.M);
                    for (uint32_t j = 0; j < 8; ++j)
                        rec[1 + j] = j < nOut ? outH[j] : 0u;
                    sHasNew[lane] = newChild;
                }
            }
            prevFull = (uint32_t)__builtin_popcountll(__ballot(full));
            __syncthreads();
            // In-level dedup, then DAG find / insert: eight lanes per item.
            bool empty = true;
            uint32_t repIdx = i8;
            if (i8 < n) {
                const uint32_t* rec = sRec[l][i8];
                empty = (rec[0] & 0xFFu) == 0u;
                for (uint32_t j = 0; j < i8 && !empty; ++j) {
                    bool eq = true;
                    for (uint32_t w = 0; w < kNodeStride && eq; ++w)
                        eq = sRec[l][j][w] == rec[w];
                    if (eq) {
                        repIdx = j;
                        break;
                    }
                }
            }
            uint32_t handle = kSentinel, created = 0;
            const bool mine = i8 < n && !empty && repIdx == i8;
            const uint8_t role = node_role(dag, (uint32_t)l);
            uint32_t r = kNotFound, hint = kNoSlab, s = 0, h = 0;
            Needle<8> nd;
            if (mine) {
                const uint32_t* rec = sRec[l][i8];
                s = node_size_from_header(rec[0]);
                needle_load(nd, rec, s, c8);
                h = needle_hash(nd, s, c8);
                // A node with a child inserted by this merge is new: no find.
                if (sHasNew[i8] == 0u)
                    r = table_find_group_hint<8>(dag.tables[s - 2], nd, s, h, c8, gbase, hint, role);
            }
            // Every find before any insert: an insert claims a slot's tag
            // before it writes the item.
            __syncthreads();
            if (mine && r == kNotFound) {
                r = table_insert_group<8>(dag.tables[s - 2], nd, s, h, c8, gbase, hint);
                created = 1u;
                if (r != kNotFound) {
                    refs_on_insert(dag, sRa, nd, s, r, role, c8);
                    if (c8 == 0)
                        atomicAdd(dag.stored, 1ull);
                }
            }
            if (mine) {
                if (r == kNotFound)
                    oom = true;
                else
                    handle = node_handle(r, s);
            }
            if (c8 == 0 && i8 < n) {
                sHandle[l][i8] = handle;
                sCreated[l][i8] = created;
            }
            __syncthreads();
            if (c8 == 0 && i8 < n && !empty && repIdx != i8) {
                sHandle[l][i8] = sHandle[l][repIdx];
                sCreated[l][i8] = sCreated[l][repIdx];
            }
            __syncthreads();
            if (__ballot(oom) != 0)
                break;
        }
        if (__ballot(oom) != 0) {
            if (lane == 0) {
                EditStatus st {};
                st.oom = 1u;
                st.rootBefore = st.rootAfter = root;
                status[k] = st;
                *ws.poison = 1u;
            }
            break;
        }
        const uint32_t newRoot = sHandle[0][0];
        if (lane == 0) {
            EditStatus st {};
            st.done = 1u;
            st.applied = 1u;
            st.seqSpine = 1u;
            st.rootBefore = root;
            st.rootAfter = newRoot;
            st.leafCount = leafCount;
            status[k] = st;
        }
        root = newRoot;
        __syncthreads();
    }
    __syncthreads();
    if (lane == 0) {
        *rootPtr = root;
        *f.waveIdx += 1u;
    }
}

// ======================================================================== host

static uint32_t ceil_log2(uint64_t v)
{
    uint32_t p = 0;
    while ((1ull << p) < v)
        ++p;
    return p;
}

// Edits whose bounds span more than a quarter of the volume on an axis run
// on their own, so that one huge edit does not pull S up for the batch.
constexpr int kMinSplit = 2;

static uint64_t max_span(const Bounds& b)
{
    uint64_t m = 0;
    for (int a = 0; a < 3; ++a)
        m = std::max<uint64_t>(m, (uint64_t)(b.hi[a] - b.lo[a] + 1));
    return m;
}

static bool fusable(const svdag_ctx* ctx, const EditorDev& e)
{
    if (!(e.kind == EK_SPHERE_ADD || e.kind == EK_SPHERE_REMOVE || e.kind == EK_BOX_ADD || e.kind == EK_BOX_REMOVE
          || e.kind == EK_PAINT))
        return false;
    return (int)ctx->dag.levels - (int)ceil_log2(max_span(edit_bounds(ctx, e))) >= kMinSplit;
}

uint32_t plan_batch(svdag_ctx* ctx, const EditorDev* eds, uint32_t first, uint32_t n, std::vector<BatchStep>& steps,
                    const std::vector<uint8_t>& alone, uint32_t waveBase)
{
    steps.clear();
    const uint32_t D = ctx->dag.levels;
    uint64_t maxSpan = 0;
    for (uint32_t k = first; k < n; ++k)
        if (fusable(ctx, eds[k]) && !alone[k])
            maxSpan = std::max<uint64_t>(maxSpan, max_span(edit_bounds(ctx, eds[k])));
    int S = maxSpan ? (int)D - (int)ceil_log2(maxSpan) : 0;
    S = std::min<int>(S, (int)ctx->dag.leafLevel);
    if (S < 1) {
        for (uint32_t k = first; k < n; ++k)
            steps.push_back({ k, k + 1, false });
        return 0;
    }
    SeedDesc* seeds = reinterpret_cast<SeedDesc*>(ctx->hFused);
    uint2* waves = reinterpret_cast<uint2*>(seeds + ctx->batchCap);
    uint32_t nWaves = waveBase;
    const uint32_t sh = D - (uint32_t)S;
    std::vector<Bounds> members;
    uint32_t waveStart = first;
    bool open = false;
    auto close = [&](uint32_t end) {
        if (!open)
            return;
        steps.push_back({ waveStart, end, true });
        waves[nWaves++] = make_uint2(waveStart, end);
        members.clear();
        open = false;
    };
    for (uint32_t k = first; k < n; ++k) {
        if (!fusable(ctx, eds[k]) || alone[k]) {
            close(k);
            steps.push_back({ k, k + 1, false });
            continue;
        }
        const Bounds b = edit_bounds(ctx, eds[k]);
        SeedDesc sd {};
        int64_t cellLo[3], cellHi[3];
        for (int a = 0; a < 3; ++a) {
            const int64_t lo = b.lo[a] >> sh, hi = b.hi[a] >> sh;
            sd.lo[a] = (uint32_t)lo;
            sd.n[a] = (uint32_t)(hi - lo + 1);
            cellLo[a] = lo << sh;
            cellHi[a] = ((hi + 1) << sh) - 1;
        }
        seeds[k] = sd;
        bool conflict = false;
        for (const Bounds& m : members) {
            bool overlap = true;
            for (int a = 0; a < 3; ++a)
                overlap &= cellLo[a] <= m.hi[a] && m.lo[a] <= cellHi[a];
            if (overlap) {
                conflict = true;
                break;
            }
        }
        if (conflict || members.size() >= kMaxWave)
            close(k);
        if (!open) {
            open = true;
            waveStart = k;
        }
        members.push_back(b);
    }
    close(n);
    return (uint32_t)S;
}

// The wave's launch sequence.  levelItems == nullptr: fixed grids (every
// kernel walks its device-side count with a grid-stride loop), as captured
// into a graph that serves any wave of the same split level.
// part: 0 the whole wave, 1 its SVO build (wave_begin .. leaf level), 2
// its merge and spine.  The build only reads items the root reaches; the
// merge, which finds and inserts, waits for mergeAfter when given.
static int enqueue_wave(svdag_ctx* ctx, uint32_t S, const uint64_t* levelItems, uint32_t seedLanes,
                        hipEvent_t mergeAfter = nullptr, int part = 0)
{
    const DagDev& dag = ctx->dag;
    const hipStream_t st = ctx->stream;
    WsDev ws = ctx->ws;
    ws.edCount = ctx->dEdCount;
    ws.edFull = ctx->dEdFull;
    const FusedDev f { ctx->dWaves, ctx->dWaveIdx, ctx->dSeeds, ctx->dSeedWork, S, ctx->dWaveIdx + 1, ctx->dSpine };
    if (part != 2) {
        wave_begin<<<1, 256, 0, st>>>(ws, f, ctx->dStatusArr);
        fused_seed<<<(seedLanes + 255) / 256, 256, 0, st>>>(ws, dag, ctx->dEditors, f, ctx->dRoot);
    }
    if (levelItems) {
        SVDAG_TRY(launch_joint_levels(ctx, ws, ctx->dEditors, S, levelItems, ctx->maxBlocks, mergeAfter, part));
    } else {
        // Grids for any wave of this split level: a wave holds at most
        // kMaxWave edits of at most 8 level-S cells each, and every item has
        // at most 8 children, so level L holds at most 8 kMaxWave 8^(L-S)
        // items (the grid-stride loops cover the counts either way).
        uint64_t bound[kMaxLevels];
        uint64_t b = 8ull * kMaxWave;
        for (uint32_t L = 0; L < kMaxLevels; ++L) {
            bound[L] = L < S ? 0ull : b;
            if (L >= S)
                b = std::min<uint64_t>(b * 8ull, 1ull << 40);
        }
        SVDAG_TRY(launch_joint_levels(ctx, ws, ctx->dEditors, S, bound, 2048, mergeAfter, part));
    }
    if (part != 1) {
        fused_spine_par<<<1, kParThreads, 0, st>>>(ws, dag, ctx->dEditors, f, ctx->dRoot, ctx->dStatusArr,
                                                   ctx->dLevelArr);
        fused_spine<<<1, 64, 0, st>>>(ws, dag, ctx->dEditors, f, ctx->dRoot, ctx->dStatusArr, ctx->dLevelArr);
    }
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

void release_wave_graphs(svdag_ctx* ctx)
{
    for (auto& g : ctx->waveGraphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphExecDestroy(g.merge);
    }
    ctx->waveGraphs.clear();
}

// One part of the wave's launch sequence captured and instantiated.
static int capture_wave_part(svdag_ctx* ctx, uint32_t S, int part, hipGraphExec_t& out)
{
    hipGraph_t graph = nullptr;
    SVDAG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_wave(ctx, S, nullptr, 8u * kMaxWave, nullptr, part);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != SVDAG_OK) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    SVDAG_HIP(e);
    const hipError_t ie = hipGraphInstantiate(&out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SVDAG_HIP(ie);
    return SVDAG_OK;
}

// The wave's graphs for split level S: its SVO build and its merge, replayed
// one after the other (so that the merge can wait for an event in between).
static int wave_graph(svdag_ctx* ctx, uint32_t S, hipGraphExec_t& build, hipGraphExec_t& merge)
{
    for (auto it = ctx->waveGraphs.begin(); it != ctx->waveGraphs.end();) {
        if (it->generation != ctx->bufferGeneration) {
            (void)hipGraphExecDestroy(it->exec);
            (void)hipGraphExecDestroy(it->merge);
            it = ctx->waveGraphs.erase(it);
            continue;
        }
        if (it->S == S && it->groupSize == ctx->groupSize) {
            build = it->exec;
            merge = it->merge;
            return SVDAG_OK;
        }
        ++it;
    }
    SVDAG_TRY(capture_wave_part(ctx, S, 1, build));
    const int rc = capture_wave_part(ctx, S, 2, merge);
    if (rc != SVDAG_OK) {
        (void)hipGraphExecDestroy(build);
        return rc;
    }
    ctx->waveGraphs.push_back({ S, ctx->groupSize, ctx->bufferGeneration, build, merge });
    return SVDAG_OK;
}

int launch_wave(svdag_ctx* ctx, const EditorDev* hEds, const BatchStep& step, uint32_t S, hipEvent_t mergeAfter)
{
    if (ctx->batchMode == SVDAG_BATCH_FUSED_GRAPH) {
        hipGraphExec_t build = nullptr, merge = nullptr;
        SVDAG_TRY(wave_graph(ctx, S, build, merge));
        SVDAG_HIP(hipGraphLaunch(build, ctx->stream));
        if (mergeAfter)
            SVDAG_HIP(hipStreamWaitEvent(ctx->stream, mergeAfter, 0));
        SVDAG_HIP(hipGraphLaunch(merge, ctx->stream));
        return SVDAG_OK;
    }
    uint64_t levelItems[kMaxLevels] {};
    for (uint32_t k = step.first; k < step.end; ++k) {
        const Bounds b = edit_bounds(ctx, hEds[k]);
        for (uint32_t L = S; L <= ctx->dag.leafLevel; ++L)
            levelItems[L] += level_bound(ctx, b, L);
    }
    return enqueue_wave(ctx, S, levelItems, 8u * (step.end - step.first), mergeAfter);
}

} // namespace svdag
