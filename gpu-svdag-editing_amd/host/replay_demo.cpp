// Plays a reference-format replay (replays/*.csv) on a terrain DAG through
// the C++ facade, the way the reference's ReplayManager drives
// Engine::edit<T> (replay.h:583-596).  Prints one line per frame:
// "<frame> <edits so far> <undo depth> <canonical hash hex>".
// --parse-only prints the parsed actions instead and touches no device.
#include "replay.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv)
{
    using namespace svdag;
    if (argc < 2) {
        std::fprintf(stderr, "usage: replay_demo <replay.csv> [levels] [replay levels] [--parse-only]\n");
        return 2;
    }
    const bool parseOnly = argc > 4 && std::strcmp(argv[4], "--parse-only") == 0;
    try {
        std::vector<ReplayAction> actions = load_replay_csv(std::string(argv[1]));
        if (parseOnly) {
            for (const ReplayAction& a : actions)
                std::printf("%d %.6g %.6g %.6g %.6g %.6g %.6g %.6g %.6g %.6g %.6g %d %u\n", (int)a.type, a.location[0],
                            a.location[1], a.location[2], a.radius, a.src[0], a.src[1], a.src[2], a.dest[0], a.dest[1],
                            a.dest[2], a.add ? 1 : 0, a.material);
            return 0;
        }
        const uint32_t levels = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 10;
        const uint32_t replayLevels = argc > 3 ? (uint32_t)std::atoi(argv[3]) : levels;
        MyGPUHashDAG dag(levels, 4);
        const double side = double(1u << levels);
        svdag_terrain t { 0x5EED, 512.0 * side / 1024.0, 200.0 * side / 1024.0, 1.0 / (256.0 * side / 1024.0), 5, 1 };
        dag.build_terrain(t);
        MyGPUHashDAGUndoRedo undoRedo;
        CameraView view;
        ReplayManager replay(std::move(actions), levels, replayLevels);
        while (!replay.at_end()) {
            replay.replay_frame(dag, undoRedo, view);
            const auto d = undoRedo.depth(dag);
            std::printf("%u %u %u %016llx\n", replay.frame(), replay.stats().edits, d[0],
                        (unsigned long long)dag.canonical_hash());
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
