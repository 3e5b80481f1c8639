// Replay driver: the reference's recorded edit traces (replays/*.csv) drive
// the MI355X back-end through the C++ facade.
//
// reference: ReplayManager::load_csv / replay_frame and the IReplayAction
// family (replay.h:15-27, 88-575, 583-683; replay.cpp:10-140).
//
// File format (as the reference reads it): one action per line,
// "<ActionName>,<field>,<field>,...", fields split on ',', trailing empty
// cells dropped.  Field counts per action are checked as the reference's
// load() functions check them.  Coordinates and radii are scaled by
// 2^sceneDepth / 2^replayDepth when applied (replay.cpp:10).
//
// What each action does here, as in the reference's GPU branch:
//   EditSphere  MyGpuSphereEditor<add>(location, radius, material = 1)
//   EditCube    MyGpuBoxEditor<add>(location, radius, material = 1)
//   EditCopy    MyGpuCopyEditor(dag, src, dest, location, radius); the CPU
//               path's transform and swirl fields are read and ignored, as
//               the reference's GPU copy ignores them
//   EditPaint   MyGpuSpherePaintEditor(location, radius, color, material = 1)
//   EditFill    no GPU counterpart (the reference applies it to the CPU
//               HashDAG only): skipped and counted
//   Undo/Redo   MyGPUHashDAGUndoRedo::undo / redo
//   SetLocation / SetRotation   the camera view
//   SetToolParameters           UI state: read and ignored
//   EndFrame    ends a frame (replay_frame returns)
#pragma once
#include "my_gpu_hash_dag.hpp"
#include <fstream>
#include <istream>
#include <string>
#include <vector>

namespace svdag {

enum class ReplayActionType {
    SetLocation,
    SetRotation,
    SetToolParameters,
    EditSphere,
    EditCube,
    EditCopy,
    EditFill,
    EditPaint,
    Undo,
    Redo,
    EndFrame,
};

struct ReplayAction {
    ReplayActionType type = ReplayActionType::EndFrame;
    float location[3] = { 0, 0, 0 };
    float src[3] = { 0, 0, 0 };
    float dest[3] = { 0, 0, 0 };
    float radius = 0.0f;
    float rotation[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
    float color[3] = { 1, 1, 1 };
    bool add = false;
    uint32_t material = 1;
};

inline ReplayActionType replay_action_from_string(const std::string& s)
{
    static const char* names[] = { "SetLocation", "SetRotation", "SetToolParameters", "EditSphere", "EditCube",
                                   "EditCopy", "EditFill", "EditPaint", "Undo", "Redo", "EndFrame" };
    for (int i = 0; i < 11; ++i)
        if (s == names[i])
            return ReplayActionType(i);
    throw Error(SVDAG_E_INVALID, "unknown replay action '" + s + "'");
}

namespace detail {
    inline void expect(bool ok, const std::string& what)
    {
        if (!ok)
            throw Error(SVDAG_E_INVALID, "replay: " + what);
    }
    inline bool parse_bool(const std::string& s)
    {
        expect(s == "true" || s == "false" || s == "TRUE" || s == "FALSE", "boolean field '" + s + "'");
        return s == "true" || s == "TRUE";
    }
    inline void parse_line(std::vector<std::string>& data, std::vector<ReplayAction>& out)
    {
        expect(!data.empty() && !data[0].empty(), "empty line");
        ReplayAction a;
        a.type = replay_action_from_string(data[0]);
        data.erase(data.begin());
        while (!data.empty() && data.back().empty())
            data.pop_back();
        const size_t n = data.size();
        auto f = [&](size_t i) { return std::stof(data[i]); };
        switch (a.type) {
        case ReplayActionType::SetLocation: // replay.h:184-190
            expect(n == 3, "SetLocation takes 3 fields");
            for (int i = 0; i < 3; ++i)
                a.location[i] = f(i);
            break;
        case ReplayActionType::SetRotation: // replay.h:217-229
            expect(n == 9, "SetRotation takes 9 fields");
            for (int i = 0; i < 9; ++i)
                a.rotation[i] = f(i);
            break;
        case ReplayActionType::SetToolParameters: // replay.h:263-283
            expect(n == 11 || n == 14, "SetToolParameters takes 11 or 14 fields");
            break;
        case ReplayActionType::EditSphere: // replay.h:328-340
        case ReplayActionType::EditCube: // replay.h:373-385
            expect(n == 5 || n == 6, "EditSphere / EditCube take 5 or 6 fields");
            for (int i = 0; i < 3; ++i)
                a.location[i] = f(i);
            a.radius = f(3);
            a.add = parse_bool(data[4]);
            if (n > 5)
                a.material = (uint32_t)std::stoi(data[5]);
            break;
        case ReplayActionType::EditCopy: // replay.h:423-447
            expect(n == 10 || n == 21, "EditCopy takes 10 or 21 fields");
            for (int i = 0; i < 3; ++i) {
                a.location[i] = f(i);
                a.src[i] = f(3 + i);
                a.dest[i] = f(6 + i);
            }
            a.radius = f(9);
            if (n == 21)
                (void)parse_bool(data[19]);
            break;
        case ReplayActionType::EditFill: // replay.h:492-498
            expect(n == 4, "EditFill takes 4 fields");
            for (int i = 0; i < 3; ++i)
                a.location[i] = f(i);
            a.radius = f(3);
            break;
        case ReplayActionType::EditPaint: // replay.h:530-542
            expect(n == 4 || n == 7 || n == 8, "EditPaint takes 4, 7 or 8 fields");
            for (int i = 0; i < 3; ++i)
                a.location[i] = f(i);
            a.radius = f(3);
            if (n > 6)
                for (int i = 0; i < 3; ++i)
                    a.color[i] = f(4 + i);
            if (n > 7)
                a.material = (uint32_t)std::stoi(data[7]);
            break;
        default:
            break;
        }
        out.push_back(a);
    }
} // namespace detail

// Reads a replay the way ReplayManager::load_csv does (replay.h:650-683); a
// UTF-8 byte-order mark and carriage returns are tolerated.
inline std::vector<ReplayAction> load_replay_csv(std::istream& in)
{
    std::vector<ReplayAction> out;
    std::vector<std::string> data;
    char c;
    bool start = true;
    while (in.get(c)) {
        if (start && (unsigned char)c == 0xEF) {
            char b1, b2;
            in.get(b1);
            in.get(b2);
            start = false;
            continue;
        }
        start = false;
        if (c == '\r')
            continue;
        if (c == '\n') {
            if (!data.empty())
                detail::parse_line(data, out);
            data.clear();
        } else if (c == ',') {
            if (data.empty())
                data.emplace_back();
            data.emplace_back();
        } else {
            if (data.empty())
                data.emplace_back();
            data.back().push_back(c);
        }
    }
    if (!data.empty())
        detail::parse_line(data, out);
    return out;
}

inline std::vector<ReplayAction> load_replay_csv(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw Error(SVDAG_E_INVALID, "replay: cannot open " + path);
    return load_replay_csv(in);
}

struct ReplayStats {
    uint32_t edits = 0;    // edits applied to the DAG
    uint32_t skipped = 0;  // EditFill (CPU-only in the reference)
    uint32_t undos = 0, redos = 0;
};

// Plays a loaded replay frame by frame (ReplayManager::replay_frame,
// replay.h:583-596).
class ReplayManager {
public:
    ReplayManager(std::vector<ReplayAction> actions, uint32_t sceneDepth, uint32_t replayDepth)
        : actions_(std::move(actions))
        , scale_((float)(1u << sceneDepth) / (float)(1u << replayDepth))
    {
    }

    bool at_end() const { return index_ >= actions_.size(); }
    uint32_t frame() const { return frame_; }
    const ReplayStats& stats() const { return stats_; }

    // Applies actions up to and including the next EndFrame.
    void replay_frame(MyGPUHashDAG& dag, MyGPUHashDAGUndoRedo& undoRedo, CameraView& view)
    {
        if (at_end())
            return;
        ++frame_;
        while (index_ < actions_.size()) {
            const ReplayAction& a = actions_[index_++];
            apply(a, dag, undoRedo, view);
            if (a.type == ReplayActionType::EndFrame)
                break;
        }
    }

private:
    float3 scaled(const float* v) const { return { v[0] * scale_, v[1] * scale_, v[2] * scale_ }; }

    void apply(const ReplayAction& a, MyGPUHashDAG& dag, MyGPUHashDAGUndoRedo& undoRedo, CameraView& view)
    {
        switch (a.type) {
        case ReplayActionType::SetLocation:
            for (int i = 0; i < 3; ++i)
                view.position[i] = a.location[i];
            break;
        case ReplayActionType::SetRotation:
            for (int i = 0; i < 9; ++i)
                view.rotation[i] = a.rotation[i];
            break;
        case ReplayActionType::EditSphere:
            if (a.add)
                editMyHashDag(MyGpuSphereEditor<true>(scaled(a.location), a.radius * scale_, a.material), dag, undoRedo);
            else
                editMyHashDag(MyGpuSphereEditor<false>(scaled(a.location), a.radius * scale_, a.material), dag, undoRedo);
            ++stats_.edits;
            break;
        case ReplayActionType::EditCube:
            if (a.add)
                editMyHashDag(MyGpuBoxEditor<true>(scaled(a.location), a.radius * scale_, a.material), dag, undoRedo);
            else
                editMyHashDag(MyGpuBoxEditor<false>(scaled(a.location), a.radius * scale_, a.material), dag, undoRedo);
            ++stats_.edits;
            break;
        case ReplayActionType::EditCopy:
            editMyHashDag(MyGpuCopyEditor<MyGPUHashDAG>(dag, scaled(a.src), scaled(a.dest), scaled(a.location),
                                                        a.radius * scale_),
                          dag, undoRedo);
            ++stats_.edits;
            break;
        case ReplayActionType::EditPaint:
            editMyHashDag(MyGpuSpherePaintEditor(scaled(a.location), a.radius * scale_,
                                                 { a.color[0], a.color[1], a.color[2] }, a.material),
                          dag, undoRedo);
            ++stats_.edits;
            break;
        case ReplayActionType::EditFill:
            ++stats_.skipped;
            break;
        case ReplayActionType::Undo:
            undoRedo.undo(dag);
            ++stats_.undos;
            break;
        case ReplayActionType::Redo:
            undoRedo.redo(dag);
            ++stats_.redos;
            break;
        default:
            break;
        }
    }

    std::vector<ReplayAction> actions_;
    float scale_;
    size_t index_ = 0;
    uint32_t frame_ = 0;
    ReplayStats stats_;
};

} // namespace svdag
