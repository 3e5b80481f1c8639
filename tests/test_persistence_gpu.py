"""Checkpoint / resume: svdag_save writes the canonical dump, svdag_load (and
svdag_import_canonical) rebuild an equal DAG in any context of the same
depth and material bits, and later edits behave identically."""
import numpy as np
import pytest

import oracle_py
import svdag_py
import workloads

L10 = 10


def _edited(levels=L10, seed=3, n=24):
    tp = workloads.terrain_params(levels)
    g = svdag_py.MyGPUHashDAG(levels, 4)
    g.build_terrain(**tp)
    o = oracle_py.OracleDag(levels, 4)
    o.build_terrain(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"], tp["material"])
    side = 1 << levels
    hts = oracle_py.terrain_heights(tp["seed"], tp["base"], tp["amplitude"], tp["frequency"], tp["octaves"],
                                    0, 0, side, side)
    edits = workloads.surface_edits(lambda x, z: hts[z, x], n, seed=seed, levels=levels, rmin=2.0, rmax=side / 10,
                                    near=side / 20)
    for e in edits:
        g.edit(*e)
        o.edit(*e)
    return g, o, hts


@pytest.mark.gpu
def test_save_load_round_trip_and_resume(tmp_path):
    a, o, hts = _edited()
    path = str(tmp_path / "dag.svdg")
    a.save(path)
    b = svdag_py.MyGPUHashDAG(L10, 4)  # empty: only the fully-filled chain
    b.load(path)
    assert b.canonical() == a.canonical()
    assert b.canonical()[0] == o.canonical()[0]
    assert b.history() == (0, 0)
    # Resume: the same further edits on the original, the loaded copy and the oracle.
    more = workloads.surface_edits(lambda x, z: hts[z, x], 12, seed=77, levels=L10, rmin=2.0, rmax=80.0, near=40.0)
    for e in more:
        a.edit(*e)
        b.edit(*e)
        o.edit(*e)
    assert a.canonical() == b.canonical()
    assert b.canonical()[0] == o.canonical()[0]
    # Loading into a context with other content replaces its root.
    c = svdag_py.MyGPUHashDAG(L10, 4)
    c.build_terrain(**workloads.terrain_params(L10))
    c.edit(svdag_py.SPHERE_ADD, 3, (500.0, 600.0, 500.0), 40.0)
    c.load(path)
    a2, _, _ = _edited()
    assert c.canonical() == a2.canonical()
    for g in (a, b, c, a2):
        g.close()


@pytest.mark.gpu
def test_import_canonical_matches_export_and_rejects_bad_input(tmp_path):
    a, _, _ = _edited(seed=5, n=10)
    h, counts, dump = a.canonical(dump=True)
    b = svdag_py.MyGPUHashDAG(L10, 4)
    b.import_canonical(dump)
    assert b.canonical() == (h, counts)
    before = b.canonical()
    bad = dump.copy()
    bad[-1] = 10 ** 7  # root id out of range
    with pytest.raises(svdag_py.SvdagError):
        b.import_canonical(bad)
    with pytest.raises(svdag_py.SvdagError):
        b.import_canonical(dump[:-5])
    path = tmp_path / "dag.svdg"
    a.save(str(path))
    raw = bytearray(path.read_bytes())
    raw[40] ^= 0x01  # inside the dump: the checksum no longer matches
    bad_path = tmp_path / "bad.svdg"
    bad_path.write_bytes(bytes(raw))
    with pytest.raises(svdag_py.SvdagError):
        b.load(str(bad_path))
    other = svdag_py.MyGPUHashDAG(11, 4)
    with pytest.raises(svdag_py.SvdagError):
        other.load(str(path))
    assert b.canonical() == before  # failed loads leave the root alone
    for g in (a, b, other):
        g.close()
