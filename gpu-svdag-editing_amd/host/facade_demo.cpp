// Drives the C++ facade the way Engine::edit<T> drives the reference
// (engine.h:104-163): build a terrain DAG, apply edits, undo/redo, collect
// garbage, trace, then write and read the reference's file format (argv[2]
// names a scene file to write).  Prints "<canonical hash hex> <hits> <lit>
// <unoccluded>" for a test to check.
#include "my_gpu_hash_dag.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv)
{
    using namespace svdag;
    const uint32_t levels = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 10;
    try {
        MyGPUHashDAG dag(levels, 4);
        const double side = double(1u << levels);
        svdag_terrain t { 0x5EED, 512.0 * side / 1024.0, 200.0 * side / 1024.0, 1.0 / (256.0 * side / 1024.0), 5, 1 };
        dag.build_terrain(t);
        MyGPUHashDAGUndoRedo undoRedo;
        const float s = (float)(side / 1024.0);
        editMyHashDag(MyGpuSphereEditor<true>({ 300 * s, 520 * s, 400 * s }, 40 * s, 2), dag, undoRedo);
        // The reference's full signature: stats sink, memory pool, stream.
        StatsRecorder stats;
        GpuMemoryPool memPool;
        editMyHashDag(MyGpuBoxEditor<false>({ 310 * s, 530 * s, 410 * s }, 20 * s, 0), dag, undoRedo, stats, memPool,
                      nullptr);
        if (stats.values.size() < 3 || stats.values[1].second != 1.0) {
            std::fprintf(stderr, "editMyHashDag reported no statistics\n");
            return 3;
        }
        editMyHashDag(MyGpuSpherePaintEditor({ 600 * s, 500 * s, 600 * s }, 50 * s, { 1, 0, 0 }, 7), dag, undoRedo);
        editMyHashDag(MyGpuCopyEditor<MyGPUHashDAG>(dag, { 300 * s, 520 * s, 400 * s }, { 700 * s, 530 * s, 200 * s },
                                                    { 300 * s, 520 * s, 400 * s }, 30 * s),
                      dag, undoRedo);
        undoRedo.undo(dag);
        undoRedo.redo(dag);
        undoRedo.garbageCollect(dag);
        // Two batches of two edits, each followed by a collection, as one
        // edit stream.
        const auto streamed = editStream(dag,
                                         { MyGpuSphereEditor<true>({ 150 * s, 560 * s, 800 * s }, 24 * s, 5).desc(),
                                           MyGpuSphereEditor<false>({ 820 * s, 480 * s, 150 * s }, 18 * s, 0).desc(),
                                           MyGpuBoxEditor<true>({ 500 * s, 600 * s, 500 * s }, 12 * s, 3).desc(),
                                           MyGpuSpherePaintEditor({ 400 * s, 510 * s, 700 * s }, 30 * s, { 0, 1, 0 }, 9)
                                               .desc() },
                                         2, true);
        if (streamed.size() != 4) {
            std::fprintf(stderr, "editStream returned %zu results\n", streamed.size());
            return 4;
        }
        DAGTracer tracer(128, 128);
        CameraView view;
        const double pos[3] = { -0.1 * side, 0.85 * side, -0.1 * side };
        for (int i = 0; i < 3; ++i)
            view.position[i] = pos[i];
        // Looking along +x+z and down (rows: -right, up, forward), as look_at().
        const double f[3] = { 0.6 * side, -0.5 * side, 0.6 * side };
        double fl = 0;
        for (double v : f)
            fl += v * v;
        fl = __builtin_sqrt(fl);
        const double fw[3] = { f[0] / fl, f[1] / fl, f[2] / fl };
        double r[3] = { fw[1] * 0 - fw[2] * 1, fw[2] * 0 - fw[0] * 0, fw[0] * 1 - fw[1] * 0 };
        double rl = __builtin_sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (double& v : r)
            v /= rl;
        const double u[3] = { r[1] * fw[2] - r[2] * fw[1], r[2] * fw[0] - r[0] * fw[2], r[0] * fw[1] - r[1] * fw[0] };
        const double rot[9] = { -r[0], -r[1], -r[2], u[0], u[1], u[2], fw[0], fw[1], fw[2] };
        for (int i = 0; i < 9; ++i)
            view.rotation[i] = rot[i];
        DAGInfo info { { 0, 0, 0 }, { side, side, side } };
        tracer.resolve_paths(view, info, dag);
        uint32_t hits = 0;
        for (const auto& h : tracer.hits())
            hits += h.material != 0xFFFFFFFFu;
        // Shadows and ambient occlusion of the traced surface, as Engine
        // resolves them after the paths (engine.cpp:743-750).
        tracer.resolve_ambient_occlusion(dag, 3, 1.0f, 32.0f);
        uint32_t open = 0;
        for (float v : tracer.ambientOcclusion())
            open += v > 0.0f;
        tracer.resolve_ambient_occlusion_blur(dag);
        tracer.resolve_shadows(dag, 1.0f);
        std::vector<float> palette;
        for (int m = 0; m < 16; ++m)
            palette.insert(palette.end(), { (float)m / 15.0f, 0.5f, 1.0f - (float)m / 15.0f });
        tracer.resolve_colors(dag, palette);
        tracer.resolve_lighting(dag, true, true, 0.0f);
        for (uint32_t c : tracer.colors())
            if ((c >> 24) != 0xFFu) {
                std::fprintf(stderr, "lighting wrote a pixel without alpha\n");
                return 6;
            }
        uint32_t lit = 0;
        for (float v : tracer.sunLight())
            lit += v > 0.0f;
        // The reference's file format: MyGPUHashDAG::writeTo / readFrom, and the
        // Engine scene file (argv[2]) with DAGInfo and the view.
        const uint64_t hash = dag.canonical_hash();
        MyGPUHashDAG copy(levels, 4);
        const std::vector<uint8_t> stream = dag.writeTo();
        if (copy.readFrom(stream.data(), stream.size()) != stream.size() || copy.canonical_hash() != hash) {
            std::fprintf(stderr, "writeTo / readFrom round trip differs\n");
            return 4;
        }
        if (argc > 2) {
            write_scene(argv[2], info, dag, view);
            MyGPUHashDAG loaded(levels, 4);
            DAGInfo info2 {};
            CameraView view2;
            read_scene(argv[2], info2, loaded, view2);
            bool same = loaded.canonical_hash() == hash;
            for (int k = 0; k < 3; ++k)
                same = same && info2.boundsAABBMax[k] == side && view2.position[k] == view.position[k];
            for (int k = 0; k < 9; ++k)
                same = same && view2.rotation[k] == view.rotation[k];
            if (!same) {
                std::fprintf(stderr, "scene file round trip differs\n");
                return 5;
            }
        }
        std::printf("%016llx %u %u %u\n", (unsigned long long)hash, hits, lit, open);
        return 0;
    } catch (const Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
