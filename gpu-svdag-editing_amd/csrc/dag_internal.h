// Internals of the editable DAG shared by dag.hip, gc.hip and tracer.hip.
#pragma once
#include "editors.h"
#include "svdag_internal.h"
#include <mutex>
#include <vector>

namespace svdag {

struct Server;
constexpr uint32_t kNodeStride = 10; // SVO node record: header + 8 children + pad

struct LevelDesc {
    uint32_t start;     // first global work index of the level
    uint32_t count;     // work items (= SVO items) of the level
    uint32_t fullCount; // nonzero iff discovery flagged a full item at this level
    uint32_t pad;
};

struct EditStatus {
    uint32_t overflow;  // workspace too small: nothing merged
    uint32_t oom;       // slab arena exhausted during the merge
    uint32_t applied;   // root updated
    uint32_t rootBefore;
    uint32_t rootAfter;
    uint32_t leafCount;
    uint32_t done;      // the edit ran to completion (applied or skipped)
    uint32_t fallback;  // the fused pass could not take this edit: run it alone
    uint32_t build;     // a DAG build (terrain, voxel grid): merged and published
                        // even without leaves (fully-filled children only)
    uint32_t seqSpine;  // a fused wave's edit whose levels above S the sequential spine built
    uint32_t pad[2];
};

// Edit workspace as seen by the kernels.
struct WsDev {
    uint4* work;                 // [capWork] (x, y, z, old handle)
    CopyState* state;            // [capWork] copy-tool parents (copy edits only)
    uint32_t* nodes;             // [capWork][kNodeStride]
    uint32_t* leaves;            // [capLeaf][leafStride]
    uint32_t* rep;               // [capWork] representative global index / kSentinel
    uint32_t* handle;            // [capWork] DAG handle of representatives
    unsigned long long* dedup;   // [dedupCap] (hash << 32 | index + 1)
    LevelDesc* lvl;              // [kMaxLevels]
    EditStatus* status;
    uint32_t capWork, capLeaf, dedupCap, leafStride;
    // Fused multi-edit pass only (null otherwise): batch-relative edit of
    // each work item, and per-(edit, level) item and full-item counters.
    uint32_t* editOf;            // [capWork]
    uint32_t* edCount;           // [batch][kMaxLevels]
    uint32_t* edFull;            // [batch][kMaxLevels]
    // 1 where the merge inserted representative w (0 where it found it); a
    // parent with such a child cannot be in the DAG yet.
    uint8_t* created;            // [capWork]
    // Where the insert of a representative the DAG find missed starts: the
    // first line of its chain the find saw with a free slot (kNoSlab: head).
    uint32_t* hint;              // [capWork]
    // Set by an edit that could not complete (overflow, arena exhausted,
    // handed back by a fused wave).  The collections of an edit stream read
    // it and release nothing once it is set; the stream clears it first.
    uint32_t* poison;
};

// Level-S cell range of an edit in a fused wave: cells lo .. lo + n - 1 per
// axis, n in {1, 2}.
struct SeedDesc {
    uint32_t lo[3];
    uint32_t n[3];
};

// Fused-wave state: the waves of a batch are launched back to back and each
// wave reads its edit range from waves[*waveIdx]; the last kernel of a wave
// advances waveIdx.
struct FusedDev {
    const uint2* waves;     // (first edit, end edit) per wave
    uint32_t* waveIdx;
    const SeedDesc* seeds;  // [batch]
    uint32_t* seedWork;     // [batch][8]: level-S work index of a cell, or kSentinel
    uint32_t S;             // split level: the fused part covers levels S .. leaf
    uint32_t* spineDone;    // set by the parallel spine when it built the wave's spine
    void* spine;            // the parallel spine's scratch (spine_scratch_bytes())
};

// Bytes of device scratch the parallel spine needs (fused.hip).
size_t spine_scratch_bytes();

// ---------------------------------------------------------------- references
// Every stored DAG item has a count of the references stored items hold to
// it (one per child word that decodes to it) and a role byte: the decodings
// under which its own child words were counted.  A node's children are nodes
// unless it sits directly above the leaf level, where they are leaves, and
// the same words can be a valid node in both places, or a leaf.  An item
// is stored with the role of the insert that made it (none for a leaf), and
// a find that returns an item of another role counts as a miss, so a new
// copy with that role is inserted: every stored item is used in one role,
// the one its children were counted under.  Roots hold no count:
// collections release the roots no longer in use and free what drops to
// zero (gc.hip).  A count that reaches kStickyRefs stays there: the most
// shared items (the full leaf, common surface patterns) are referenced
// millions of times, and atomics on one word serialise, so a cached load in
// front of every update skips them.  Such an item is kept until a marking
// collection.  64: on a depth-15 terrain edit stream the stream ran 7%
// faster than at 2048, and a marking collection afterwards found nothing
// the counting ones had kept.
constexpr uint8_t kRoleParentOfNodes = 1, kRoleParentOfLeaves = 2;
constexpr uint32_t kStickyRefs = 64;
__device__ inline uint8_t node_role(const DagDev& dag, uint32_t L)
{
    return L + 1u == dag.leafLevel ? kRoleParentOfLeaves : kRoleParentOfNodes;
}
// The tables' count and role arrays, staged where a kernel reads them often
// (LDS): one dependent load less per count or role access.
constexpr uint32_t kRefTables = 20; // item sizes 2 .. 18 words use 17
struct RefArrays {
    uint32_t* refs[kRefTables];
    uint8_t* roles[kRefTables];
};
__device__ inline void ref_arrays_stage(RefArrays& r, const DagDev& dag)
{
    if (threadIdx.x < dag.numTables) {
        r.refs[threadIdx.x] = dag.tables[threadIdx.x].refs;
        r.roles[threadIdx.x] = dag.tables[threadIdx.x].roles;
    }
}
__device__ inline void child_ref_add(const DagDev& dag, const RefArrays& ra, uint32_t h, uint8_t role)
{
    const bool leaf = role == kRoleParentOfLeaves;
    const uint32_t s = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
    const uint32_t ptr = leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h);
    uint32_t* c = &ra.refs[s - 2][ptr];
    if (ld_filter(c) < kStickyRefs)
        atomicAdd(c, 1u);
}
__device__ inline void child_ref_add(const DagDev& dag, uint32_t h, uint8_t role)
{
    const bool leaf = role == kRoleParentOfLeaves;
    const uint32_t s = leaf ? leaf_handle_size(dag, h) : node_handle_size(h);
    const uint32_t ptr = leaf ? leaf_handle_ptr(dag, h) : node_handle_ptr(h);
    uint32_t* c = &dag.tables[s - 2].refs[ptr];
    if (ld_filter(c) < kStickyRefs)
        atomicAdd(c, 1u);
}
// node handle h may be used where its children decode under `role`.
__device__ inline bool node_has_role(const DagDev& dag, uint32_t h, uint8_t role)
{
    return (dag.tables[node_handle_size(h) - 2].roles[node_handle_ptr(h)] & role) != 0;
}
// leaf handle h names an item stored as a leaf (no roles).
__device__ inline bool leaf_is_plain(const DagDev& dag, uint32_t h)
{
    return dag.tables[leaf_handle_size(dag, h) - 2].roles[leaf_handle_ptr(dag, h)] == 0;
}
// After a group inserted the item held in n (s words) at pointer res of
// table s - 2: record its role (0 for a leaf) and count its children.
template <int G, int MW>
__device__ inline void refs_on_insert(const DagDev& dag, const RefArrays& ra, const Needle<G, MW>& n, uint32_t s,
                                      uint32_t res, uint8_t role, uint32_t gl)
{
    if (gl == 0)
        ra.roles[s - 2][res] = role;
    if (role) {
#pragma unroll
        for (int k = 0; k < Needle<G, MW>::K; ++k) {
            const uint32_t j = gl + (uint32_t)k * G;
            if (j >= 1u && j < s)
                child_ref_add(dag, ra, n.w[k], role);
        }
    }
}

// A leaf is fully filled iff all 64 voxels are set with one material
// (my_gpu_hash_dag_factory.cpp:357-360).
__device__ inline bool leaf_is_full(const uint32_t* leaf, uint32_t M, uint32_t& mat)
{
    if (leaf[0] != 0xFFFFFFFFu || leaf[1] != 0xFFFFFFFFu)
        return false;
    if (M == 0) {
        mat = 0;
        return true;
    }
    const uint32_t m = leaf[2] & ((1u << M) - 1u);
    const uint32_t rep = (M == 4) ? m * 0x11111111u : m * 0x01010101u;
    const uint32_t words = (64u * M) / 32u;
    for (uint32_t i = 0; i < words; ++i)
        if (leaf[2 + i] != rep)
            return false;
    mat = m;
    return true;
}

} // namespace svdag

struct svdag_ctx {
    svdag_config cfg {};
    svdag::TableSet ts;
    svdag::DagDev dag {};
    uint32_t* dFullyFilled = nullptr;
    uint32_t* dRoot = nullptr;
    uint32_t root = svdag::kSentinel;
    svdag::WsDev ws {};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evWait = nullptr; // svdag_edit_on_stream: the caller's stream position
    std::vector<uint32_t> undo, redo;
    uint32_t groupSize = 16;
    uint32_t maxBlocks = 8192; // grid cap of the per-level launches (autotuned)
    std::vector<uint64_t> slabLimit; // per table: most lines it may hold (table_line_limit)
    // Device address of this GPU's copy of the file-scope query staging array
    // (svdag_query_voxels_noalloc), taken once per device at create.
    uint32_t* queryScratch = nullptr;
    // Host-visible status block (pinned): EditStatus followed by the level table.
    svdag::EditStatus* hStatus = nullptr;
    svdag::LevelDesc* hLevels = nullptr;
    // Terrain height pyramid.
    int32_t* hmin = nullptr;
    int32_t* hmax = nullptr;
    uint64_t pyrOffset[svdag::kMaxLevels] {};
    // Garbage-collection frontier buffers.
    uint32_t* gcFront[2] = { nullptr, nullptr };
    uint64_t gcCap = 0;
    uint32_t* gcCounters = nullptr;
    uint32_t gcCtrIdx = 0; // the counter block the next counting collection uses
    // Per-slot visit bytes of every table (kept between collections and
    // regrown with the arenas) and their device pointer array.
    std::vector<uint8_t*> gcMasks;
    std::vector<uint32_t> gcMaskCap;
    uint8_t** dGcMasks = nullptr;
    // Per-edit device slots of a batch (editors, statuses, level tables) and
    // their pinned host mirror.
    svdag::EditorDev* dEditors = nullptr;
    svdag::EditStatus* dStatusArr = nullptr;
    svdag::LevelDesc* dLevelArr = nullptr;
    void* hBatch = nullptr;
    uint32_t batchCap = 0;
    // Fused-pass slots (sized with the batch slots).
    uint32_t* dEdCount = nullptr;
    uint32_t* dEdFull = nullptr;
    svdag::SeedDesc* dSeeds = nullptr;
    uint32_t* dSeedWork = nullptr;
    uint2* dWaves = nullptr;
    uint32_t* dWaveIdx = nullptr;
    void* hFused = nullptr;     // pinned: SeedDesc[batchCap] then uint2[batchCap]
    uint32_t batchMode = 1;     // 0: edits one after another, 1: fused waves, 2: fused waves as HIP graphs
    // Captured wave graphs, keyed by (split level, lanes per item, buffer
    // generation); the generation changes whenever a buffer a graph points
    // at is reallocated.
    struct WaveGraph {
        uint32_t S, groupSize;
        uint64_t generation;
        hipGraphExec_t exec;  // the wave's SVO build (from wave_begin)
        hipGraphExec_t merge; // its merge and spine, replayed after exec
    };
    std::vector<WaveGraph> waveGraphs;
    uint64_t bufferGeneration = 0;
    // Tracer output buffer.
    void* dHits = nullptr;
    size_t hitCap = 0;
    // Surface codes of the last primary trace (0: no surface, else
    // 1 + 2 * axis + (normal points along +axis)), its size and root, for
    // the shadow and ambient-occlusion passes.
    uint8_t* dSurf = nullptr;
    float* dShade = nullptr; // per-pixel sun visibility of the last shadow pass (hitCap floats)
    float* dAo = nullptr;    // per-pixel result of the last ambient-occlusion pass (hitCap floats)
    uint32_t* dColor = nullptr; // per-pixel diffuse colour (RGBA8) of the last colour pass
    // Which of those hold results for the current primary trace, and that
    // trace's camera (position, first ray, per-pixel steps) for the lighting
    // pass's fog.
    bool sunValid = false, aoValid = false, colorValid = false;
    double traceCam[12] = {};
    // Path-tracing accumulation (3 floats per pixel) and what it was
    // accumulated for.
    float* dAccum = nullptr;
    size_t accumCap = 0;
    bool accumValid = false;
    uint32_t numAccumulated = 0, accumW = 0, accumH = 0, accumRoot = 0xFFFFFFFFu, accumDepth = 0, accumIntegrate = 0;
    float accumEnv = 0.0f;
    double accumCam[12] = {};
    void* dSpine = nullptr;  // parallel spine scratch, allocated with the batch slots
    // Reference counting (gc.hip): items stored in the tables, every root the
    // context has held since the last collection, and whether the counts
    // may be off (an edit rolled back after inserting, an import), in which
    // case the next collection marks from the roots and recounts.
    unsigned long long* dStored = nullptr;
    std::vector<uint32_t> rootLog;
    bool refsDirty = false;
    // Live items after the last marking collection (or the first one) and
    // after the last collection of either kind.
    unsigned long long liveAtMark = 0, liveAfterGc = 0;
    uint32_t* dGcCand = nullptr;         // released roots of a collection
    uint32_t gcCandCap = 0;
    uint2* dGcTouchedList = nullptr;     // (table, bucket) pairs for the chain pass
    std::vector<uint32_t*> gcTouched;    // per table: a bit per bucket
    uint32_t traceW = 0, traceH = 0, traceRoot = 0xFFFFFFFFu;
    // Collections run so far: between two collections no slot is freed, so
    // a root handle names one DAG content.
    uint64_t gcSerial = 0;
    // Fused-wave edits whose spine the one-wave sequential kernel built (the
    // parallel spine declined the wave), and edits a wave handed back to run
    // on their own, since the context was made.
    uint64_t seqSpineEdits = 0, handedBackEdits = 0;
    // Items per released root in each release round of recent counting
    // collections, a decaying maximum (sizes the next collection's grids).
    std::vector<double> gcPerRoot;
    // A counting collection nobody waited for (svdag_gc without stats): its
    // counters land in pinned hGcCtr when evGc completes, and gc_settle
    // takes them in before anything reads the collection's results.
    bool gcPending = false;
    uint32_t gcPendingRounds = 0, gcPendingCand = 0;
    void* hGcCtr = nullptr;
    hipEvent_t evGc = nullptr;
    // An edit stream's results beside the edit statuses: its last
    // collection's counters and its "a collection stopped short" word
    // (pinned), and that word on the device.
    void* hStream = nullptr;
    uint32_t* dStreamFail = nullptr;
    // An edit stream runs each batch's collection on gcStream beside the
    // next batch's SVO build; per batch, one event marks the batch's end on
    // the context's stream and one the collection's end on gcStream.
    hipStream_t gcStream = nullptr;
    std::vector<hipEvent_t> streamEvents;
    // The reference-format stream a size query built (refformat.hip), kept
    // for the call that fetches it while (root, gcSerial) is unchanged.
    std::vector<uint8_t> refStream;
    uint32_t refStreamRoot = 0xFFFFFFFFu;
    uint64_t refStreamGc = ~0ull;
    std::mutex mu;
    // Serving (serve.hip).  While a server runs, frame traces read the tables
    // on frameStream; whoever reallocates or sweeps tables first takes
    // frameMu and drains that stream (FrameQuiesce).
    svdag::Server* server = nullptr;
    std::mutex frameMu;
    hipStream_t frameStream = nullptr;
    // The last frame of a stopped server.
    void* dLastFrame = nullptr;
    uint32_t lastFrameW = 0, lastFrameH = 0, lastFrameRoot = svdag::kSentinel;
};

namespace svdag {
// Holds frame traces off the tables while they are reallocated or swept.
struct FrameQuiesce {
    std::unique_lock<std::mutex> lock;
    explicit FrameQuiesce(svdag_ctx* ctx) : lock(ctx->frameMu)
    {
        if (ctx->frameStream)
            (void)hipStreamSynchronize(ctx->frameStream);
    }
};
int ctx_refresh_tables(svdag_ctx* ctx);
// Every root the context holds is logged when it is replaced: the next
// collection releases the logged roots (gc.hip).
inline void root_log(svdag_ctx* ctx)
{
    if (ctx->root != 0xFFFFFFFFu)
        ctx->rootLog.push_back(ctx->root);
}
// mark: force the marking collection (else reference counts, when valid).
int gc_run(svdag_ctx* ctx, svdag_gc_stats* st, bool mark = false, bool wait = true);
int gc_settle(svdag_ctx* ctx);
int gc_stream_prepare(svdag_ctx* ctx);
int gc_enqueue_stream(svdag_ctx* ctx, uint32_t first, uint32_t count, const std::vector<uint32_t>& held,
                      uint32_t* fail, void** ctrOut, hipStream_t st);
void gc_stream_take(svdag_ctx* ctx, const void* counters, uint32_t roots);
size_t gc_counters_bytes();
// surfDev (optional) receives one surface code per pixel (see dSurf).
int launch_trace(svdag_ctx* ctx, const svdag_camera& cam, uint32_t W, uint32_t H, uint32_t root, hipStream_t stream,
                 svdag_hit* outDev, uint8_t* surfDev = nullptr);
// Run edits [0, n) in order on ctx->stream (one sync at the end), with the
// overflow / OOM retry; used by svdag_edit, svdag_edit_batch and the server.
int run_edits(svdag_ctx* ctx, const EditorDev* eds, uint32_t n, svdag_edit_stats* stats, float* totalMs);
int make_editor(const svdag_ctx* ctx, const svdag_edit_desc& d, EditorDev& e);

// Canonical dump (export.hip): for level = leaf level down to 0,
// [level, count, then per item size, words...] with children as indices into
// the level below's list, then the root's index.  canonical_dump writes the
// reachable DAG with each level's unique contents sorted; import_dump takes
// any dump of that shape whose items are unique within each level, finds or
// inserts them, and makes its root the context's root (the caller holds
// ctx->mu).
int canonical_dump(svdag_ctx* ctx, std::vector<uint32_t>& dump, std::vector<uint32_t>& counts);
int import_dump(svdag_ctx* ctx, const uint32_t* d, size_t len);
uint64_t fnv1a(const std::vector<uint32_t>& words);

// Voxel bounds [lo, hi] (inclusive) outside which an edit changes nothing.
struct Bounds {
    int64_t lo[3], hi[3];
};
Bounds edit_bounds(const svdag_ctx* ctx, const EditorDev& e);
uint64_t level_bound(const svdag_ctx* ctx, const Bounds& b, uint32_t L);

// One edit's kernels on ctx->stream (dag.hip).
int launch_edit(svdag_ctx* ctx, const EditorDev& ed, const EditorDev* edDev, const WsDev& ws);
// Levels S .. leaf of a fused wave: SVO build and merge with per-item editors
// (dag.hip).  levelItems[L] bounds the items of level L (grid sizing only).
// part: 0 the whole sequence, 1 the SVO build only, 2 the merge only.
int launch_joint_levels(svdag_ctx* ctx, const WsDev& ws, const EditorDev* eds, uint32_t S, const uint64_t* levelItems,
                        uint32_t gridCap = 8192, hipEvent_t mergeAfter = nullptr, int part = 0);

// Fused multi-edit waves (fused.hip).
struct BatchStep {
    uint32_t first, end;  // edits [first, end)
    bool fused;           // a fused wave (else a single edit on its own)
};
// Splits edits [first, n) into steps; fills the seeds / waves host mirrors
// (the waves from index waveBase on) and returns the split level S (0 when
// nothing can be fused).  Edits flagged in `alone` run on their own.
uint32_t plan_batch(svdag_ctx* ctx, const EditorDev* eds, uint32_t first, uint32_t n, std::vector<BatchStep>& steps,
                    const std::vector<uint8_t>& alone, uint32_t waveBase = 0);
// mergeAfter: the wave's merge kernels wait for this event (its SVO build
// does not).
int launch_wave(svdag_ctx* ctx, const EditorDev* hEds, const BatchStep& step, uint32_t S,
                hipEvent_t mergeAfter = nullptr);
void release_wave_graphs(svdag_ctx* ctx);
}
