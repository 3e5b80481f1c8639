"""Replays drive the GPU back-end exactly as they drive the oracle: the C++
player (host/replay.hpp via replay_demo) and the Python player over
svdag_py both give the oracle's canonical DAG after every frame, with undo
and redo."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py
import replay
import svdag_py
import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(os.path.dirname(HERE), "gpu-svdag-editing_amd", "lib", "replay_demo")
LEVELS, REPLAY_LEVELS = 10, 11  # scale 0.5, as a replay recorded one level deeper


def _synthetic_replay(seed=12):
    rng = np.random.default_rng(seed)
    tp = workloads.terrain_params(LEVELS)
    side = 1 << LEVELS
    hts = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                    0, 0, side, side)
    acts = []
    for frame in range(12):
        acts.append({"type": "SetLocation", "location": [float(v) for v in rng.integers(0, 2 * side, 3)]})
        acts.append({"type": "SetRotation", "rotation": [1, 0, 0, 0, 1, 0, 0, 0, 1]})
        for _ in range(int(rng.integers(1, 4))):
            x, z = (int(v) for v in rng.integers(8, side - 8, 2))
            # replay space is twice the scene: halves stay exact in float32
            loc = [2.0 * x, 2.0 * float(hts[z, x]) + float(rng.integers(-40, 40)), 2.0 * z]
            r = float(rng.integers(4, 120)) / 2.0
            k = int(rng.integers(0, 5))
            if k == 0:
                acts.append({"type": "EditSphere", "location": loc, "radius": r, "add": True,
                             "material": int(rng.integers(1, 16))})
            elif k == 1:
                acts.append({"type": "EditSphere", "location": loc, "radius": r, "add": False, "material": 1})
            elif k == 2:
                acts.append({"type": "EditCube", "location": loc, "radius": r, "add": bool(rng.integers(0, 2)),
                             "material": int(rng.integers(1, 16))})
            elif k == 3:
                acts.append({"type": "EditPaint", "location": loc, "radius": r, "material": int(rng.integers(1, 16))})
            else:
                d = [loc[0] + float(rng.integers(-200, 200)), loc[1] + 20.0, loc[2] + float(rng.integers(-200, 200))]
                acts.append({"type": "EditCopy", "location": loc, "src": loc, "dest": d, "radius": r,
                             "full": bool(rng.integers(0, 2))})
        if frame in (4, 9):
            acts.append({"type": "Undo"})
            acts.append({"type": "Undo"})
            acts.append({"type": "Redo"})
        if frame == 6:
            acts.append({"type": "EditFill", "location": [100.0, 100.0, 100.0], "radius": 10.0})
        acts.append({"type": "EndFrame"})
    return acts


def _oracle_frames(acts):
    tp = workloads.terrain_params(LEVELS)
    o = oracle_py.OracleDag(LEVELS, 4)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    frames = []
    replay.play(acts, o, LEVELS, REPLAY_LEVELS, history=replay.RootHistory(),
                on_frame=lambda f: frames.append(o.canonical()[0]))
    return frames


@pytest.mark.gpu
def test_cpp_replay_player_matches_oracle(tmp_path):
    acts = replay.load_replay_csv(replay.write_replay_csv(_synthetic_replay()))
    path = tmp_path / "synthetic.csv"
    path.write_text(replay.write_replay_csv(_synthetic_replay()))
    r = subprocess.run([DEMO, str(path), str(LEVELS), str(REPLAY_LEVELS)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.strip().split("\n")]
    got = [int(row[3], 16) for row in rows]
    want = _oracle_frames(acts)
    assert len(got) == len(want) == 12
    assert got == want


@pytest.mark.gpu
def test_python_replay_player_on_gpu_matches_oracle():
    acts = replay.load_replay_csv(replay.write_replay_csv(_synthetic_replay(seed=13)))
    tp = workloads.terrain_params(LEVELS)
    g = svdag_py.MyGPUHashDAG(LEVELS, 4)
    g.build_terrain(**tp)
    frames = []
    edits, skipped = replay.play(acts, g, LEVELS, REPLAY_LEVELS, on_frame=lambda f: frames.append(g.canonical()[0]))
    assert skipped == 1 and edits > 12
    assert frames == _oracle_frames(acts)
    g.close()
