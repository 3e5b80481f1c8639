"""The GPU factory build from a dense voxel grid against the oracle's factory
rule (my_gpu_hash_dag_factory.cpp:317-424) and against the grid itself."""
import numpy as np
import pytest

import oracle_py
import svdag_py

NONE = 0xFFFFFFFF


def _grid(side, rng, materials):
    vox = np.full((side, side, side), 0xFF, np.uint8)
    c = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).astype(np.float64)
    for _ in range(6):
        ctr = rng.uniform(0, side, 3)
        r = rng.uniform(side / 10, side / 4)
        vox[np.sum((c + 0.5 - ctr) ** 2, -1) <= r * r] = rng.choice(materials)
    h = side // 4
    vox[:h, :h, :h] = materials[0]                       # a uniform block: filled nodes above the leaves
    vox[h:2 * h, :h, :h] = 0xFF                          # and an empty one
    vox[side - 4:, side - 4:, side - 4:] = materials[-1]  # one full leaf at the far corner
    return vox


def _voxels(g, side):
    xyz = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.uint32)
    return g.query(xyz).reshape(side, side, side)


@pytest.mark.gpu
@pytest.mark.parametrize("levels,grid_log2,material_bits", [(7, 7, 4), (7, 5, 4), (6, 6, 8), (6, 6, 0)])
def test_voxel_build_matches_factory_rule_and_grid(levels, grid_log2, material_bits):
    rng = np.random.default_rng(levels * 10 + grid_log2 + material_bits)
    mats = [0] if material_bits == 0 else list(range(1, min(15, (1 << material_bits) - 2)))
    g = svdag_py.MyGPUHashDAG(levels, material_bits)
    gs = 1 << grid_log2
    vox = _grid(gs, rng, mats)
    g.build_voxels(vox)
    full = np.full((1 << levels,) * 3, 0xFF, np.uint8)
    full[:gs, :gs, :gs] = vox
    o = oracle_py.OracleDag(levels, material_bits)
    oracle_py.build_dense(o, full)
    assert g.canonical()[0] == o.canonical()[0]
    want = np.where(full == 0xFF, NONE, full.astype(np.uint32))
    assert np.array_equal(_voxels(g, 1 << levels), want)
    assert g.history() == (0, 0)
    g.close()


@pytest.mark.gpu
def test_voxel_build_edge_cases():
    g = svdag_py.MyGPUHashDAG(6, 4)
    # empty grid: empty DAG; uniform grid: the fully-filled chain's node
    g.build_voxels(np.full((64, 64, 64), 0xFF, np.uint8))
    assert g.root == NONE
    g.build_voxels(np.full((64, 64, 64), 5, np.uint8))
    o = oracle_py.OracleDag(6, 4)
    oracle_py.build_dense(o, np.full((64, 64, 64), 5, np.uint8))
    assert g.canonical()[0] == o.canonical()[0]
    # a material the 4 material bits cannot hold is rejected, the DAG kept
    bad = np.full((16, 16, 16), 0xFF, np.uint8)
    bad[3, 4, 5] = 16
    with pytest.raises(svdag_py.SvdagError):
        g.build_voxels(bad)
    assert g.canonical()[0] == o.canonical()[0]
    # larger than the DAG is rejected
    with pytest.raises(svdag_py.SvdagError):
        g.build_voxels(np.zeros((128, 128, 128), np.uint8))
    # edits after a voxel build behave as after any other build
    g.edit(svdag_py.SPHERE_REMOVE, 0, (32.0, 32.0, 32.0), 10.0)
    o.edit(svdag_py.SPHERE_REMOVE, 0, (32.0, 32.0, 32.0), 10.0)
    assert g.canonical()[0] == o.canonical()[0]
    g.close()


LEAFLESS_GRIDS = ["octant", "half_spaces", "uniform_small", "octant_tree"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", LEAFLESS_GRIDS)
def test_voxel_build_without_leaves(case):
    # Grids whose every aligned 4^3 cell is uniform or empty make no leaf in
    # the edit pipeline.  The reference drops such an edit
    # (my_gpu_hash_dag_edits.cu:651), but a build is not an edit: the result
    # must still be the factory rule's DAG (fully-filled nodes only).
    print(f"[case] leafless voxel build {case}", flush=True)
    levels = 7
    side = 1 << levels
    if case == "octant":
        gs, vox = side, np.full((side,) * 3, 0xFF, np.uint8)
        vox[:64, 64:, :64] = 6
    elif case == "half_spaces":
        gs, vox = side, np.full((side,) * 3, 3, np.uint8)
        vox[:, 64:, :] = 9
    elif case == "uniform_small":
        gs, vox = 32, np.full((32,) * 3, 4, np.uint8)
    else:
        gs, vox = side, np.full((side,) * 3, 0xFF, np.uint8)
        vox[:32, :32, :32] = 2
        vox[96:, 32:64, 64:96] = 11
        vox[64:80, :16, 16:32] = 5
    g = svdag_py.MyGPUHashDAG(levels, 4)
    g.build_voxels(vox)
    full = np.full((side,) * 3, 0xFF, np.uint8)
    full[:gs, :gs, :gs] = vox
    o = oracle_py.OracleDag(levels, 4)
    oracle_py.build_dense(o, full)
    assert o.root != NONE
    assert g.root != NONE
    assert g.canonical()[0] == o.canonical()[0]
    rng = np.random.default_rng(3)
    pts = rng.integers(0, side, size=(4000, 3)).astype(np.uint32)
    want = full[pts[:, 0], pts[:, 1], pts[:, 2]].astype(np.uint32)
    want[want == 0xFF] = NONE
    assert np.array_equal(g.query(pts), want)
    # edits still skip when they make no leaf
    before = g.canonical()[0]
    st = g.edit(svdag_py.BOX_ADD, 7, (32.0, 96.0, 32.0), 32.0)
    assert st.applied == 0 and g.canonical()[0] == before
    g.close()
