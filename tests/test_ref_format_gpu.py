"""The library's reader and writer of the reference's version-3 file format
(svdag_read/write_reference_dag, svdag_read/write_reference_scene) against
the oracle's restatement (oracle/ref_format.py):

  - the library writes the same bytes as the restatement for the same DAG
    (material bits 0, 4 and 8), and reads them back to the same canonical DAG
    in a fresh context;
  - a stream the restatement wrote reads into the same DAG, and edits after
    that equal the oracle's (the flow of the reference's "Paint inside epic
    citadel" case, gpu_hash_dag_tests.cpp:264-320: load a scene, paint,
    compare voxel by voxel);
  - a scene file with voxel textures is read past its textures to its end;
  - damaged streams and mismatched contexts are rejected with an error.
"""
import struct

import numpy as np
import pytest

import oracle_py
import ref_format
import svdag_py
import workloads


def _pair(levels, M):
    g = svdag_py.MyGPUHashDAG(levels, M)
    o = oracle_py.OracleDag(levels, M)
    tp = dict(workloads.terrain_params(levels))
    tp["material"] = min(tp["material"], (1 << M) - 1) if M else 0
    g.build_terrain(**tp)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    side = float(1 << levels)
    edits = [(svdag_py.SPHERE_ADD, 3, (side * 0.3, side * 0.55, side * 0.4), side * 0.12),
             (svdag_py.SPHERE_REMOVE, 0, (side * 0.6, side * 0.45, side * 0.5), side * 0.1),
             (svdag_py.BOX_ADD, 6, (side * 0.7, side * 0.6, side * 0.2), side * 0.05)]
    if M:
        edits.append((svdag_py.PAINT, 5, (side * 0.5, side * 0.45, side * 0.5), side * 0.2))
    for kind, mat, c, r in edits:
        mat = mat % (1 << M) if M else 0
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    assert g.canonical()[0] == o.canonical()[0]
    return g, o


@pytest.mark.gpu
@pytest.mark.parametrize("M", [0, 4, 8])
def test_written_stream_equals_the_restatement_and_reads_back(M):
    levels = 9
    g, o = _pair(levels, M)
    h, counts, dump = o.canonical(dump=True)
    data = g.write_reference_dag()
    want = ref_format.write_dag(levels, M, dump)
    assert len(data) == len(want)
    assert data == want
    # The restatement's reader sees the oracle DAG in the library's bytes.
    lv, m2, _root, recs, used, _ff = ref_format.read_dag(data)
    assert (lv, m2, used) == (levels, M, len(data))
    assert oracle_py.canonical_from_records(recs)[0] == h
    # And the library reads them into a fresh context.
    g2 = svdag_py.MyGPUHashDAG(levels, M)
    assert g2.read_reference_dag(data + b"trailing") == len(data)
    assert g2.canonical() == (h, counts)
    g2.close()
    g.close()


@pytest.mark.gpu
def test_scene_from_the_restatement_loads_and_edits_like_the_oracle(tmp_path):
    levels, M = 10, 4
    g0, o = _pair(levels, M)
    g0.close()
    h, _, dump = o.canonical(dump=True)
    img = (np.arange(4 * 5 * 3, dtype=np.uint8), 5, 3)
    mats = [("rock", [img, img, img], [0.1] * 9)]
    bounds = (-10.0, -20.0, -30.0, 40.0, 50.0, 60.0)
    view = tuple(0.5 * i for i in range(12))
    path = tmp_path / "scene.gpu_hash_dag.dag.bin"
    path.write_bytes(ref_format.write_scene(levels, M, dump, bounds, view, mats))
    g = svdag_py.MyGPUHashDAG(levels, M)
    b2, v2 = g.read_reference_scene(str(path))
    assert b2 == bounds and v2 == view
    assert g.canonical()[0] == h
    # Paint and a sphere after the load, as the reference's loaded-scene case.
    for kind, mat, c, r in [(svdag_py.PAINT, 3, (612.0, 430.0, 505.0), 15.0),
                            (svdag_py.SPHERE_ADD, 7, (300.0, 520.0, 700.0), 40.0)]:
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    assert g.canonical()[0] == o.canonical()[0]
    pts = np.random.default_rng(5).integers(0, 1 << levels, size=(200000, 3)).astype(np.uint32)
    assert np.array_equal(g.query(pts), o.query(pts))
    g.close()


@pytest.mark.gpu
def test_scene_written_by_the_library_round_trips(tmp_path):
    levels, M = 9, 8
    g, o = _pair(levels, M)
    path = str(tmp_path / "out.dag.bin")
    g.write_reference_scene(path, bounds=(1, 2, 3, 4, 5, 6), view=tuple(range(12)))
    b, recs, v, names = ref_format.read_scene(open(path, "rb").read())
    assert b == (1, 2, 3, 4, 5, 6) and v == tuple(float(x) for x in range(12)) and names == []
    assert oracle_py.canonical_from_records(recs)[0] == o.canonical()[0]
    g2 = svdag_py.MyGPUHashDAG(levels, M)
    b2, v2 = g2.read_reference_scene(path)
    assert b2 == (1, 2, 3, 4, 5, 6) and v2 == tuple(float(x) for x in range(12))
    assert g2.canonical() == g.canonical()
    # Defaults: [0, 2^depth]^3 bounds and CameraView's identity rotation.
    g.write_reference_scene(path)
    b3, v3 = g2.read_reference_scene(path)
    assert b3 == (0, 0, 0, 512, 512, 512) and v3 == (0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1)
    g2.close()
    g.close()


@pytest.mark.gpu
def test_damaged_streams_and_other_contexts_are_rejected(tmp_path):
    levels, M = 8, 4
    g, o = _pair(levels, M)
    data = g.write_reference_dag()
    before = g.canonical()
    # The root field follows the version, fullyFilledNodes and its width/height.
    root_off = 4 + 12 + 4 * levels * (1 << M) + 8
    assert struct.unpack_from("<I", data, root_off)[0] != 0xFFFFFFFF
    bad_cases = {
        "truncated": data[:len(data) // 2],
        "version 2": struct.pack("<i", 2) + data[4:],
        # A node handle whose overflow slab lies beyond every super block.
        "dangling root": data[:root_off] + struct.pack("<I", (6 << 29) | ((1 << 22) << 6)) + data[root_off + 4:],
    }
    for name, bad in bad_cases.items():
        with pytest.raises(svdag_py.SvdagError) as e:
            g.read_reference_dag(bad)
        assert e.value.code == -1, name  # SVDAG_E_INVALID
    # A failed read leaves the context's DAG as it was.
    assert g.canonical() == before
    for lv, m in [(levels + 1, M), (levels, 8)]:
        other = svdag_py.MyGPUHashDAG(lv, m)
        with pytest.raises(svdag_py.SvdagError):
            other.read_reference_dag(data)
        other.close()
    with pytest.raises(svdag_py.SvdagError):
        g.read_reference_scene(str(tmp_path / "missing.bin"))
    # A scene whose textures run past the end of the file.
    path = tmp_path / "short.bin"
    g.write_reference_scene(str(path))
    raw = path.read_bytes()
    path.write_bytes(raw[:-8] + struct.pack("<Q", 3))
    with pytest.raises(svdag_py.SvdagError):
        g.read_reference_scene(str(path))
    g.close()
    assert o.canonical()[0] == before[0]
