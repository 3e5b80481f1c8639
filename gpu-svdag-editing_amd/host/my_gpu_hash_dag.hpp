// C++ host facade over the C ABI (include/svdag.h) with the reference's
// interface names, so a caller written against EDag::MyGpuDag reads the same:
//
//   reference                                          here
//   MyGPUHashDAG<EMemoryType::GPU_Malloc>              svdag::MyGPUHashDAG
//     (my_gpu_hash_dag.h:59-339)
//   MyGpuSphereEditor<bool>, MyGpuBoxEditor<bool>,     same names
//   MyGpuSpherePaintEditor, MyGpuCopyEditor<TDAG>
//     (my_gpu_dag_editors.h:82-253, 457-536)
//   editMyHashDag(editor, dag, undoRedo, stats,        svdag::editMyHashDag(editor, dag, undoRedo)
//     memPool, stream) (my_gpu_hash_dag_edits.h:12-13)
//   MyGPUHashDAGUndoRedo::{undo,redo,garbageCollect}   svdag::MyGPUHashDAGUndoRedo
//     (my_gpu_base_dag.h:16-82)
//   DAGTracer::resolve_paths(camera, dagInfo, dag)     svdag::DAGTracer::resolve_paths
//     (dag_tracer.cu:112-134)
//
// Ownership differs in one respect, by design: the DAG object owns its
// device memory (RAII) where the reference needs an explicit free().  The
// undo/redo history lives in the DAG context (the reference keeps it in a
// separate MyGPUHashDAGUndoRedo object that swaps the root); the facade's
// MyGPUHashDAGUndoRedo forwards to it.  Errors throw svdag::Error (the
// reference aborts on CUDA errors and traps on failed checks).
#pragma once
#include "svdag.h"
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace svdag {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m)
        : std::runtime_error("svdag error " + std::to_string(c) + ": " + m)
        , code(c)
    {
    }
};

inline void check(int code)
{
    if (code != SVDAG_OK)
        throw Error(code, svdag_last_error());
}

struct float3 {
    float x, y, z;
};

class MyGPUHashDAG {
public:
    static constexpr uint32_t invalid_handle = 0xFFFFFFFFu;

    explicit MyGPUHashDAG(uint32_t levels, uint32_t materialBits = 4, uint32_t loadFactor = 75)
    {
        svdag_config cfg {};
        cfg.levels = levels;
        cfg.material_bits = materialBits;
        cfg.load_factor = loadFactor;
        check(svdag_create(&cfg, &ctx_));
        levels_ = levels;
    }
    MyGPUHashDAG(const MyGPUHashDAG&) = delete;
    MyGPUHashDAG& operator=(const MyGPUHashDAG&) = delete;
    MyGPUHashDAG(MyGPUHashDAG&& o) noexcept
        : ctx_(o.ctx_)
        , levels_(o.levels_)
    {
        o.ctx_ = nullptr;
    }
    ~MyGPUHashDAG() { free(); }
    // MyGPUHashDAG::free (my_gpu_hash_dag.cu:68-76): releases the device
    // memory now; the destructor does it otherwise.  Idempotent.
    void free()
    {
        if (ctx_)
            svdag_destroy(ctx_);
        ctx_ = nullptr;
    }

    uint32_t levels() const { return levels_; }
    uint32_t leaf_level() const { return levels_ - 2; }
    uint32_t get_first_node_index() const
    {
        uint32_t r = 0;
        check(svdag_root(ctx_, &r));
        return r;
    }

    void build_terrain(const svdag_terrain& t) { check(svdag_build_terrain(ctx_, &t)); }
    // The factory build (my_gpu_hash_dag_factory.cpp:317-508) from a dense grid
    // materials[(x * side + y) * side + z] of side 2^sideLog2, 0xFF = empty.
    void build_from_voxels(const uint8_t* materials, uint32_t sideLog2)
    {
        check(svdag_build_voxels(ctx_, materials, sideLog2));
    }

    // DAGUtils::get_value over many voxels: 0xFFFFFFFF marks an empty voxel.
    std::vector<uint32_t> get_values(const std::vector<std::array<uint32_t, 3>>& xyz) const
    {
        std::vector<uint32_t> out(xyz.size());
        check(svdag_query_voxels(ctx_, reinterpret_cast<const uint32_t*>(xyz.data()), (uint32_t)xyz.size(), out.data()));
        return out;
    }

    uint64_t canonical_hash() const
    {
        uint64_t h = 0;
        check(svdag_export_canonical(ctx_, &h, nullptr, nullptr, nullptr));
        return h;
    }

    // MyGPUHashDAG::writeTo / readFrom (my_gpu_hash_dag.cu:157-177): the
    // reference's version-3 stream of the reachable DAG.  readFrom takes one
    // stream from the front of data and returns the bytes it used.
    std::vector<uint8_t> writeTo() const
    {
        size_t n = 0;
        check(svdag_write_reference_dag(ctx_, nullptr, &n));
        std::vector<uint8_t> out(n);
        check(svdag_write_reference_dag(ctx_, out.data(), &n));
        out.resize(n);
        return out;
    }
    size_t readFrom(const uint8_t* data, size_t len)
    {
        size_t used = 0;
        check(svdag_read_reference_dag(ctx_, data, len, &used));
        return used;
    }

    svdag_ctx* context() const { return ctx_; }

private:
    svdag_ctx* ctx_ = nullptr;
    uint32_t levels_ = 0;
};

// ------------------------------------------------------------------ editors
template <bool isAdding>
struct MyGpuSphereEditor {
    float3 center;
    float radius;
    uint32_t fillMaterial;
    MyGpuSphereEditor(float3 c, float r, uint32_t material)
        : center(c)
        , radius(r)
        , fillMaterial(material)
    {
    }
    svdag_edit_desc desc() const
    {
        return { isAdding ? (uint32_t)SVDAG_SPHERE_ADD : (uint32_t)SVDAG_SPHERE_REMOVE, fillMaterial,
                 { center.x, center.y, center.z }, radius, { 0, 0, 0 }, { 0, 0, 0 }, 0u };
    }
};

template <bool isAdding>
struct MyGpuBoxEditor {
    float3 center;
    float radius;
    uint32_t fillMaterial;
    MyGpuBoxEditor(float3 c, float r, uint32_t material)
        : center(c)
        , radius(r)
        , fillMaterial(material)
    {
    }
    svdag_edit_desc desc() const
    {
        return { isAdding ? (uint32_t)SVDAG_BOX_ADD : (uint32_t)SVDAG_BOX_REMOVE, fillMaterial,
                 { center.x, center.y, center.z }, radius, { 0, 0, 0 }, { 0, 0, 0 }, 0u };
    }
};

struct MyGpuSpherePaintEditor {
    float3 center;
    float radius;
    float3 paintColor; // carried for interface parity; shading is not part of this back-end
    uint32_t paintMaterial;
    MyGpuSpherePaintEditor(float3 c, float r, float3 color, uint32_t material)
        : center(c)
        , radius(r)
        , paintColor(color)
        , paintMaterial(material)
    {
    }
    svdag_edit_desc desc() const
    {
        return { (uint32_t)SVDAG_PAINT, paintMaterial, { center.x, center.y, center.z }, radius, { 0, 0, 0 }, { 0, 0, 0 },
                 0u };
    }
};

template <typename TDag>
struct MyGpuCopyEditor {
    float3 source, dest, center;
    float radius;
    uint32_t quarterTurns;
    // The reference copies the DAG value into the editor to read the source;
    // here the source is the same live DAG the edit is applied to.  The
    // optional quarter turns about +Y have no reference counterpart (the
    // reference's copy only translates).
    MyGpuCopyEditor(const TDag&, float3 src, float3 dst, float3 c, float r, uint32_t turns = 0)
        : source(src)
        , dest(dst)
        , center(c)
        , radius(r)
        , quarterTurns(turns)
    {
    }
    svdag_edit_desc desc() const
    {
        return { (uint32_t)SVDAG_COPY, 0u, { center.x, center.y, center.z }, radius, { source.x, source.y, source.z },
                 { dest.x, dest.y, dest.z }, quarterTurns };
    }
};

// ------------------------------------------------------------------ undo / redo / GC
class MyGPUHashDAGUndoRedo {
public:
    struct Frame {
        uint32_t firstNodeIndex = 0;
    };
    // MyGPUHashDAGUndoRedo::add_frame (my_gpu_base_dag.h:23-33): the frame's
    // root goes on the undo stack (not again if it is the last frame there)
    // and the redo stack is cleared.  The history lives in the DAG's context,
    // so the DAG is named.  Edits add their own frames, as editMyHashDag does.
    void add_frame(MyGPUHashDAG& dag, const Frame& frame)
    {
        check(svdag_add_frame(dag.context(), frame.firstNodeIndex));
    }
    void undo(MyGPUHashDAG& dag) { check(svdag_undo(dag.context())); }
    void redo(MyGPUHashDAG& dag) { check(svdag_redo(dag.context())); }
    // Clears the history, then frees every stored item the root and the
    // fully-filled chain do not reach (by reference counts; svdag_gc).
    svdag_gc_stats garbageCollect(MyGPUHashDAG& dag)
    {
        svdag_gc_stats st {};
        check(svdag_gc(dag.context(), &st));
        return st;
    }
    std::array<uint32_t, 2> depth(const MyGPUHashDAG& dag) const
    {
        std::array<uint32_t, 2> out {};
        check(svdag_history(dag.context(), out.data()));
        return out;
    }
};

// Applies one edit and records the previous root as an undo frame.
template <typename Editor>
svdag_edit_stats editMyHashDag(const Editor& editor, MyGPUHashDAG& dag, MyGPUHashDAGUndoRedo&)
{
    const svdag_edit_desc d = editor.desc();
    svdag_edit_stats st {};
    check(svdag_edit(dag.context(), &d, &st));
    return st;
}

// An edit stream (svdag_edit_stream; no reference counterpart): the
// reference's loop of editMyHashDag and garbageCollect, over batches of
// `batch` edits (editor.desc() of any tool), enqueued back to back with one
// wait at the end.  collectEachBatch: a collection after every batch (the
// history is then cleared); otherwise each edit leaves its undo frame.
inline std::vector<svdag_edit_stats> editStream(MyGPUHashDAG& dag, const std::vector<svdag_edit_desc>& edits,
                                                uint32_t batch, bool collectEachBatch, float* gpuMs = nullptr)
{
    std::vector<svdag_edit_stats> st(edits.size());
    check(svdag_edit_stream(dag.context(), edits.data(), (uint32_t)edits.size(), batch, collectEachBatch ? 1u : 0u,
                            st.data(), gpuMs));
    return st;
}

// The reference's statistics sink (stats.h): named values reported per edit.
struct StatsRecorder {
    std::vector<std::pair<std::string, double>> values;
    void report(const std::string& name, double value) { values.emplace_back(name, value); }
};
// The reference's transient device allocator (memory.h:192-224).  The context
// allocates its edit workspace once and keeps it, so there is nothing to pass.
struct GpuMemoryPool {
};

// The reference's full signature (my_gpu_hash_dag_edits.h:12-13): the edit
// starts after the work queued on `stream` (a hipStream_t, or nullptr) and has
// finished when the call returns; its GPU time, edit-SVO items per level and
// whether it was applied are reported to `stats`.
template <typename Editor, typename Stream>
svdag_edit_stats editMyHashDag(const Editor& editor, MyGPUHashDAG& dag, MyGPUHashDAGUndoRedo&, StatsRecorder& stats,
                               GpuMemoryPool&, Stream stream)
{
    const svdag_edit_desc d = editor.desc();
    svdag_edit_stats st {};
    check(svdag_edit_on_stream(dag.context(), &d, &st, (void*)stream));
    stats.report("editing", st.gpu_ms);
    stats.report("applied", st.applied);
    for (uint32_t l = 0; l < 32 && st.svo_items[l]; ++l)
        stats.report("edit SVO items, level " + std::to_string(l), st.svo_items[l]);
    return st;
}

// ------------------------------------------------------------------ tracer
struct CameraView {
    double position[3] = { 0, 0, 0 };
    double rotation[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 }; // rows: -right, up, forward
    double fov = 60.0;
};
struct DAGInfo {
    double boundsAABBMin[3];
    double boundsAABBMax[3];
};

// Engine::readFrom / writeTo (engine.cpp:1543-1557): the scene file the
// reference loads at start-up, for its DAG, DAGInfo and camera view (voxel
// textures are skipped on read and written empty).
inline void read_scene(const std::string& path, DAGInfo& info, MyGPUHashDAG& dag, CameraView& view)
{
    double b[6], v[12];
    check(svdag_read_reference_scene(dag.context(), path.c_str(), b, v));
    for (int k = 0; k < 3; ++k) {
        info.boundsAABBMin[k] = b[k];
        info.boundsAABBMax[k] = b[3 + k];
        view.position[k] = v[k];
    }
    for (int k = 0; k < 9; ++k)
        view.rotation[k] = v[3 + k];
}
inline void write_scene(const std::string& path, const DAGInfo& info, const MyGPUHashDAG& dag, const CameraView& view)
{
    double b[6], v[12];
    for (int k = 0; k < 3; ++k) {
        b[k] = info.boundsAABBMin[k];
        b[3 + k] = info.boundsAABBMax[k];
        v[k] = view.position[k];
    }
    for (int k = 0; k < 9; ++k)
        v[3 + k] = view.rotation[k];
    check(svdag_write_reference_scene(dag.context(), path.c_str(), b, v));
}

class DAGTracer {
public:
    DAGTracer(uint32_t width, uint32_t height)
        : width_(width)
        , height_(height)
        , hits_((size_t)width * height)
    {
    }
    // Returns the device time in milliseconds; hit records land in hits().
    float resolve_paths(const CameraView& view, const DAGInfo& info, const MyGPUHashDAG& dag)
    {
        svdag_camera cam {};
        for (int i = 0; i < 3; ++i) {
            cam.position[i] = view.position[i];
            cam.bounds_min[i] = info.boundsAABBMin[i];
            cam.bounds_max[i] = info.boundsAABBMax[i];
        }
        for (int i = 0; i < 9; ++i)
            cam.rotation[i] = view.rotation[i];
        cam.fov_degrees = view.fov;
        float ms = 0.0f;
        check(svdag_trace_primary(dag.context(), &cam, width_, height_, hits_.data(), &ms));
        return ms;
    }
    const std::vector<svdag_hit>& hits() const { return hits_; }

    // DAGTracer::resolve_shadows (dag_tracer.cu:198-221) over the surface of
    // the last resolve_paths; sun visibility * max(N.L, 0) lands in sunLight().
    float resolve_shadows(const MyGPUHashDAG& dag, float shadowBias)
    {
        sun_.resize(hits_.size());
        float ms = 0.0f;
        check(svdag_trace_shadows(dag.context(), shadowBias, sun_.data(), &ms));
        return ms;
    }
    // DAGTracer::resolve_ambient_occlusion (dag_tracer.cu:163-187); the
    // occlusion seed advances per call as the reference's does.
    float resolve_ambient_occlusion(const MyGPUHashDAG& dag, uint32_t numSamples, float shadowBias, float aoRayLength)
    {
        ao_.resize(hits_.size());
        float ms = 0.0f;
        check(svdag_trace_ambient_occlusion(dag.context(), numSamples, shadowBias, aoRayLength, aoSeed_++, ao_.data(),
                                            &ms));
        return ms;
    }
    // DAGTracer::resolve_ambient_occlusion_blur (dag_tracer.cu:189-196): the
    // reference's 9-tap Gaussian over ambientOcclusion(), in place.
    float resolve_ambient_occlusion_blur(const MyGPUHashDAG& dag)
    {
        ao_.resize(hits_.size());
        float ms = 0.0f;
        check(svdag_blur_ambient_occlusion(dag.context(), ao_.data(), &ms));
        return ms;
    }
    // DAGTracer::resolve_colors (dag_tracer.cu:137-161) with one colour per
    // material (rgb[3 m .. 3 m + 2]) in place of the voxel textures.
    float resolve_colors(const MyGPUHashDAG& dag, const std::vector<float>& materialRgb)
    {
        colors_.resize(hits_.size());
        float ms = 0.0f;
        check(svdag_trace_colors(dag.context(), materialRgb.data(), (uint32_t)(materialRgb.size() / 3), colors_.data(),
                                 &ms));
        return ms;
    }
    // DAGTracer::resolve_lighting (dag_tracer.cu:223-254): the final RGBA8
    // image in colors(); the reference's call leaves sunBrightness at 0.
    float resolve_lighting(const MyGPUHashDAG& dag, bool applyAmbientOcclusion, bool applyShadows, float fogDensity,
                           float sunBrightness = 0.0f)
    {
        final_.resize(hits_.size());
        const svdag_lighting p { applyAmbientOcclusion ? 1u : 0u, applyShadows ? 1u : 0u, fogDensity, sunBrightness };
        float ms = 0.0f;
        check(svdag_trace_lighting(dag.context(), &p, final_.data(), &ms));
        return ms;
    }
    // DAGTracer::resolve_path_tracing (dag_tracer.cu:256-313): the final RGBA8
    // image in colors(); returns the samples accumulated so far (the
    // reference's return value).  The random seed advances per call from
    // 123, as the reference's shared occlusion / path seed does.
    uint32_t resolve_path_tracing(const MyGPUHashDAG& dag, const std::vector<float>& materialRgb,
                                  float environmentBrightness, uint32_t numSamples, uint32_t maxPathDepth,
                                  bool integratePixel)
    {
        final_.resize(hits_.size());
        const svdag_path_tracing p { environmentBrightness, numSamples, maxPathDepth, integratePixel ? 1u : 0u,
                                     aoSeed_++, 1u, 0u };
        uint32_t accumulated = 0;
        float ms = 0.0f;
        check(svdag_trace_path_tracing(dag.context(), &p, materialRgb.data(), (uint32_t)(materialRgb.size() / 3),
                                       final_.data(), &accumulated, &ms));
        return accumulated;
    }
    const std::vector<float>& sunLight() const { return sun_; }
    const std::vector<float>& ambientOcclusion() const { return ao_; }
    const std::vector<uint32_t>& diffuseColors() const { return colors_; }
    const std::vector<uint32_t>& colors() const { return final_; }

private:
    uint32_t width_, height_;
    std::vector<svdag_hit> hits_;
    std::vector<float> sun_, ao_;
    std::vector<uint32_t> colors_, final_;
    uint64_t aoSeed_ = 123; // the reference starts its occlusion seed at 123 (dag_tracer.h:66)
};

} // namespace svdag
