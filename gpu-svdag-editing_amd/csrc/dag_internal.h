// Internals of the editable DAG shared by dag.hip, gc.hip and tracer.hip.
#pragma once
#include "editors.h"
#include "svdag_internal.h"
#include <mutex>
#include <vector>

namespace svdag {

constexpr uint32_t kNodeStride = 10; // SVO node record: header + 8 children + pad

struct LevelDesc {
    uint32_t start;     // first global work index of the level
    uint32_t count;     // work items (= SVO items) of the level
    uint32_t fullCount; // items flagged full by discovery at this level
    uint32_t pad;
};

struct EditStatus {
    uint32_t overflow;  // workspace too small: nothing merged
    uint32_t oom;       // slab arena exhausted during the merge
    uint32_t applied;   // root updated
    uint32_t rootBefore;
    uint32_t rootAfter;
    uint32_t leafCount;
    uint32_t pad0, pad1;
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
};

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
    std::vector<uint32_t> undo, redo;
    uint32_t groupSize = 16;
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
    // Per-edit device slots of a batch (editors, statuses, level tables) and
    // their pinned host mirror.
    svdag::EditorDev* dEditors = nullptr;
    svdag::EditStatus* dStatusArr = nullptr;
    svdag::LevelDesc* dLevelArr = nullptr;
    void* hBatch = nullptr;
    uint32_t batchCap = 0;
    // Tracer output buffer.
    void* dHits = nullptr;
    size_t hitCap = 0;
    std::mutex mu;
};

namespace svdag {
int ctx_refresh_tables(svdag_ctx* ctx);
int launch_trace(svdag_ctx* ctx, const svdag_camera& cam, uint32_t W, uint32_t H, uint32_t root, hipStream_t stream,
                 svdag_hit* outDev);
// Run edits [0, n) in order on ctx->stream (one sync at the end), with the
// overflow / OOM retry; used by svdag_edit, svdag_edit_batch and the server.
int run_edits(svdag_ctx* ctx, const EditorDev* eds, uint32_t n, svdag_edit_stats* stats, float* totalMs);
int make_editor(const svdag_ctx* ctx, const svdag_edit_desc& d, EditorDev& e);
}
