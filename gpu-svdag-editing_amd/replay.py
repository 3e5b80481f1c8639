"""Reader and player for the reference's replay files (replays/*.csv).

Python twin of host/replay.hpp (same parsing rules, same action mapping), so a
replay can drive both the GPU back-end (svdag_py.MyGPUHashDAG) and any other
target with the same edit / undo / redo interface.

reference: ReplayManager::load_csv (replay.h:650-683), the IReplayAction
load() field checks (replay.h:184-542) and the GPU branch of apply()
(replay.cpp:10-140): coordinates and radii are scaled in float32 by
2^scene_depth / 2^replay_depth; EditFill has no GPU counterpart and is
skipped; the copy's transform and swirl fields are ignored by the GPU copy.
"""
import numpy as np

SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, BOX_REMOVE, PAINT, COPY = range(6)

ACTIONS = ("SetLocation", "SetRotation", "SetToolParameters", "EditSphere", "EditCube", "EditCopy", "EditFill",
           "EditPaint", "Undo", "Redo", "EndFrame")

_FIELDS = {"SetLocation": (3,), "SetRotation": (9,), "SetToolParameters": (11, 14), "EditSphere": (5, 6),
           "EditCube": (5, 6), "EditCopy": (10, 21), "EditFill": (4,), "EditPaint": (4, 7, 8), "Undo": (0,),
           "Redo": (0,), "EndFrame": (0,)}


def _bool(s):
    if s not in ("true", "false", "TRUE", "FALSE"):
        raise ValueError(f"replay: boolean field {s!r}")
    return s in ("true", "TRUE")


def _parse_line(fields):
    name = fields[0]
    if name not in ACTIONS:
        raise ValueError(f"replay: unknown action {name!r}")
    data = list(fields[1:])
    while data and data[-1] == "":
        data.pop()
    if len(data) not in _FIELDS[name]:
        raise ValueError(f"replay: {name} takes {_FIELDS[name]} fields, got {len(data)}")
    def f(i, n=1):
        v = [np.float32(x) for x in data[i:i + n]]
        return v if n > 1 else v[0]

    a = {"type": name}
    if name == "SetLocation":
        a["location"] = f(0, 3)
    elif name == "SetRotation":
        a["rotation"] = f(0, 9)
    elif name in ("EditSphere", "EditCube"):
        a.update(location=f(0, 3), radius=f(3), add=_bool(data[4]), material=int(data[5]) if len(data) > 5 else 1)
    elif name == "EditCopy":
        a.update(location=f(0, 3), src=f(3, 3), dest=f(6, 3), radius=f(9))
        if len(data) == 21:
            _bool(data[19])
    elif name == "EditFill":
        a.update(location=f(0, 3), radius=f(3))
    elif name == "EditPaint":
        a.update(location=f(0, 3), radius=f(3), material=int(data[7]) if len(data) > 7 else 1)
    return a


def load_replay_csv(text):
    """Actions of a replay file's text (a UTF-8 BOM and CRs are tolerated)."""
    if text.startswith("\ufeff"):
        text = text[1:]
    out = []
    for line in text.replace("\r", "").split("\n"):
        if line == "":
            continue
        out.append(_parse_line(line.split(",")))
    return out


class RootHistory:
    """Undo/redo over a target that exposes a settable `root`
    (MyGPUHashDAGUndoRedo semantics, my_gpu_base_dag.h:16-64)."""

    def __init__(self):
        self.undo_frames, self.redo_frames = [], []

    def add_frame(self, root):
        if self.undo_frames and self.undo_frames[-1] == root:
            return
        self.redo_frames.clear()
        self.undo_frames.append(root)

    def undo(self, target):
        if self.undo_frames:
            self.redo_frames.append(target.root)
            target.root = self.undo_frames.pop()

    def redo(self, target):
        if self.redo_frames:
            self.undo_frames.append(target.root)
            target.root = self.redo_frames.pop()


def play(actions, target, scene_depth, replay_depth, history=None, on_frame=None):
    """Apply actions to `target` (edit(kind, material, center, radius, src, dst)).
    Undo/redo use target.undo()/redo() unless a RootHistory is given, in which
    case it records a frame before every edit.  on_frame(frame_index) runs
    after every EndFrame.  Returns (edits applied, fills skipped)."""
    scale = np.float32(float(1 << scene_depth) / float(1 << replay_depth))
    edits = skipped = frame = 0

    def sc(v):
        return tuple(float(np.float32(x) * scale) for x in v)

    def edit(kind, mat, c, r, src=(0.0, 0.0, 0.0), dst=(0.0, 0.0, 0.0)):
        if history is not None:
            history.add_frame(target.root)
        target.edit(kind, mat, c, r, src, dst)

    for a in actions:
        t = a["type"]
        if t in ("EditSphere", "EditCube"):
            kind = {("EditSphere", True): SPHERE_ADD, ("EditSphere", False): SPHERE_REMOVE,
                    ("EditCube", True): BOX_ADD, ("EditCube", False): BOX_REMOVE}[(t, a["add"])]
            edit(kind, a["material"], sc(a["location"]), float(np.float32(a["radius"]) * scale))
            edits += 1
        elif t == "EditCopy":
            edit(COPY, 0, sc(a["location"]), float(np.float32(a["radius"]) * scale), sc(a["src"]), sc(a["dest"]))
            edits += 1
        elif t == "EditPaint":
            edit(PAINT, a["material"], sc(a["location"]), float(np.float32(a["radius"]) * scale))
            edits += 1
        elif t == "EditFill":
            skipped += 1
        elif t == "Undo":
            history.undo(target) if history is not None else target.undo()
        elif t == "Redo":
            history.redo(target) if history is not None else target.redo()
        elif t == "EndFrame":
            frame += 1
            if on_frame is not None:
                on_frame(frame)
    return edits, skipped


def write_replay_csv(actions):
    """Text of a replay in the reference's format (ReplayManager::write_csv,
    replay.h:619-632: name, then every field, comma separated)."""
    def fmt(x):
        return f"{float(x):.6f}"

    lines = []
    for a in actions:
        t = a["type"]
        if t == "SetLocation":
            fields = [fmt(x) for x in a["location"]]
        elif t == "SetRotation":
            fields = [fmt(x) for x in a["rotation"]]
        elif t in ("EditSphere", "EditCube"):
            fields = [fmt(x) for x in a["location"]] + [fmt(a["radius"]), "true" if a["add"] else "false",
                                                         str(a.get("material", 1))]
        elif t == "EditCopy":
            fields = ([fmt(x) for x in a["location"]] + [fmt(x) for x in a["src"]] + [fmt(x) for x in a["dest"]]
                      + [fmt(a["radius"])])
            if a.get("full", False):
                fields += [fmt(x) for x in (1, 0, 0, 0, 1, 0, 0, 0, 1)] + ["false", fmt(1.0)]
        elif t == "EditFill":
            fields = [fmt(x) for x in a["location"]] + [fmt(a["radius"])]
        elif t == "EditPaint":
            fields = [fmt(x) for x in a["location"]] + [fmt(a["radius"]), fmt(1), fmt(0), fmt(0),
                                                         str(a.get("material", 1))]
        else:
            fields = []
        lines.append(",".join([t] + fields))
    return "\n".join(lines) + "\n"
