"""Replay readers (host/replay.hpp and replay.py) on lines of the reference's
own recorded replays (tests/golden/replay_excerpt.csv, made by
tests/golden/make_replay_excerpt.py), following ReplayManager::load_csv and
the IReplayAction load() checks (replay.h:184-683)."""
import os
import subprocess

import numpy as np
import pytest

import replay

HERE = os.path.dirname(os.path.abspath(__file__))
EXCERPT = os.path.join(HERE, "golden", "replay_excerpt.csv")
DEMO = os.path.join(os.path.dirname(HERE), "gpu-svdag-editing_amd", "lib", "replay_demo")


def test_python_reader_on_reference_lines():
    acts = replay.load_replay_csv(open(EXCERPT).read())
    types = [a["type"] for a in acts]
    assert types.count("EditSphere") == 3 and types.count("EditCopy") == 2 and types.count("EndFrame") == 3
    s = acts[2]
    assert s["type"] == "EditSphere" and s["add"] and s["material"] == 1 and s["radius"] == np.float32(3)
    assert [float(v) for v in s["location"]] == [14483.0, 33332.0, 24591.0]
    # Trailing empty cells are dropped (SetLocation with 10 cells, EndFrame with 10 empty cells).
    loc = acts[6]
    assert loc["type"] == "SetLocation" and float(loc["location"][0]) == float(np.float32(-956.389906))
    cp = acts[7]
    assert [float(v) for v in cp["src"]] == [15983.0, 32860.0, 25406.0] and cp["radius"] == np.float32(1)
    paints = [a for a in acts if a["type"] == "EditPaint"]
    assert paints[0]["material"] == 1 and paints[1]["material"] == 6
    assert types[-3:] == ["Undo", "Redo", "EndFrame"]


@pytest.mark.parametrize("bad", ["EditSphere,1,2,3,4", "EditSphere,1,2,3,4,maybe", "EditCopy,1,2,3", "Jump,1"])
def test_python_reader_rejects_malformed_lines(bad):
    with pytest.raises(ValueError):
        replay.load_replay_csv(bad + "\n")


def test_writer_round_trip():
    acts = [{"type": "SetLocation", "location": [1.5, 2.0, -3.25]},
            {"type": "EditSphere", "location": [10, 20, 30], "radius": 4.5, "add": True, "material": 3},
            {"type": "EditCube", "location": [11, 21, 31], "radius": 2.0, "add": False, "material": 1},
            {"type": "EditCopy", "location": [1, 2, 3], "src": [4, 5, 6], "dest": [7, 8, 9], "radius": 5.0,
             "full": True},
            {"type": "EditPaint", "location": [5, 6, 7], "radius": 8.0, "material": 9},
            {"type": "Undo"}, {"type": "Redo"}, {"type": "EndFrame"}]
    back = replay.load_replay_csv(replay.write_replay_csv(acts))
    assert [a["type"] for a in back] == [a["type"] for a in acts]
    assert back[1]["material"] == 3 and back[1]["add"] and not back[2]["add"]
    assert [float(v) for v in back[3]["dest"]] == [7.0, 8.0, 9.0]
    assert back[4]["material"] == 9


def _py_lines(acts):
    kinds = list(replay.ACTIONS)
    out = []
    for a in acts:
        loc = a.get("location", [0, 0, 0]) if a["type"] not in ("SetRotation",) else [0, 0, 0]
        src = a.get("src", [0, 0, 0])
        dst = a.get("dest", [0, 0, 0])
        r = a.get("radius", 0)
        out.append("%d %.6g %.6g %.6g %.6g %.6g %.6g %.6g %.6g %.6g %.6g %d %u" % (
            kinds.index(a["type"]), *[float(np.float32(v)) for v in loc], float(np.float32(r)),
            *[float(np.float32(v)) for v in src], *[float(np.float32(v)) for v in dst],
            1 if a.get("add", False) else 0, a.get("material", 1)))
    return out


def test_cpp_reader_matches_python_reader():
    if not os.path.exists(DEMO):
        pytest.skip("replay_demo is not built (gpu-svdag-editing_amd/build_lib.py)")
    r = subprocess.run([DEMO, EXCERPT, "10", "10", "--parse-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = r.stdout.strip().split("\n")
    want = _py_lines(replay.load_replay_csv(open(EXCERPT).read()))
    assert got == want


def test_cpp_reader_rejects_malformed_file(tmp_path):
    if not os.path.exists(DEMO):
        pytest.skip("replay_demo is not built (gpu-svdag-editing_amd/build_lib.py)")
    p = tmp_path / "bad.csv"
    p.write_text("EditSphere,1,2,3,4\n")
    r = subprocess.run([DEMO, str(p), "10", "10", "--parse-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "EditSphere" in r.stderr
