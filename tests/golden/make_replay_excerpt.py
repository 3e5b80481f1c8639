"""Writes tests/golden/replay_excerpt.csv: a few lines of the reference's
recorded replays (data files under replays/ in mathijs727/GPU-SVDAG-Editing),
one or more per action type, to pin the replay readers' parsing against real
files.  Run from the repository root with the reference checked out at
/root/reference; the output is committed so tests do not need the reference.
"""
import os

REF = "/root/reference/replays"
PICKS = [  # (file, action, how many lines of that action)
    ("epiccitadel_edits.csv", "SetLocation", 1),
    ("epiccitadel_edits.csv", "SetRotation", 1),
    ("epiccitadel_edits.csv", "EditSphere", 3),
    ("epiccitadel_edits.csv", "EndFrame", 1),
    ("epiccitadel_copy.csv", "SetLocation", 1),
    ("epiccitadel_copy.csv", "EditCopy", 2),
    ("epiccitadel_copy.csv", "EndFrame", 1),
    ("epiccitadel_video_paint.csv", "SetToolParameters", 1),
    ("epiccitadel_video_paint.csv", "EditPaint", 1),
    ("epiccitadel_top_paint_pg24.csv", "EditPaint", 1),
    ("epiccitadel_video_barrel_fill.csv", "EditFill", 1),
    ("epiccitadel_video_large_carving_undo.csv", "Undo", 1),
    ("epiccitadel_video_large_carving_undo.csv", "Redo", 1),
    ("epiccitadel_video_large_carving_undo.csv", "EndFrame", 1),
]


def main():
    out = []
    for name, action, n in PICKS:
        text = open(os.path.join(REF, name), encoding="utf-8-sig").read().replace("\r", "")
        got = [l for l in text.split("\n") if l.split(",")[0] == action][:n]
        assert len(got) == n, (name, action)
        out += got
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "replay_excerpt.csv"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
