"""Pins the oracle's restatement of the reference editing path (CPU only):
the reference's own encoding tests and KATs, brute-force voxelisation, the
factory full-flag rule, the merge quirks, and a brute-force ray march."""
import numpy as np
import pytest

import oracle_py
import workloads

SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, BOX_REMOVE, PAINT, COPY = range(6)
NONE = 0xFFFFFFFF


def morton3d(x, y, z):
    out = 0
    for i in range(21):
        out |= ((x >> i) & 1) << (3 * i + 2) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i)
    return out


def test_morton_kats():
    # utils_tests.cu:6-28.
    assert morton3d(0b11, 0b11, 0b11) == 0b111111
    assert morton3d(0b00, 0b11, 0b11) == 0b011011
    assert morton3d(0b11, 0b00, 0b11) == 0b101101
    assert morton3d(0b11, 0b11, 0b00) == 0b110110
    assert morton3d(0b100000000000, 0b100000000000, 0b100000000000) & 0xFFFFFFFF == 0o7 << 33 & 0xFFFFFFFF
    for i in range(21):
        assert morton3d(1 << i, 0, 0) == 1 << (3 * i + 2)
        assert morton3d(0, 1 << i, 0) == 1 << (3 * i + 1)
        assert morton3d(0, 0, 1 << i) == 1 << (3 * i)


@pytest.mark.parametrize("payload", [1, 2, 3, 4, 79, 90, 127, 255])
@pytest.mark.parametrize("mask", [1, 2, 3, 4, 79, 90, 127, 255])
def test_header_round_trip(payload, mask):
    # gpu_base_dag_test.cpp:5-21.
    h = oracle_py.header_make(4, mask, payload)
    m, p, full, _ = oracle_py.header_decode(4, h)
    assert (m, p, full) == (mask, payload, 0)
    assert h & 0xFF == mask


@pytest.mark.parametrize("bits", [4, 8])
def test_leaf_builder_fixed_value(bits):
    # gpu_base_dag_test.cpp:24-60.
    for mat in range((1 << bits) - 1):
        for mask in (0x1, 0xFFFFFFFFFFFFFFFF, 0xAAAAAAAAAAAAAAAA, 0xABABABABABABABAB):
            present = [(mask >> i) & 1 for i in range(64)]
            leaf = oracle_py.leaf_build(bits, present, [mat] * 64)
            assert oracle_py.leaf_size(bits, leaf) == len(leaf)
            padded = np.concatenate([leaf, np.full(20 - len(leaf), 0xFFFFFFFF, np.uint32)])
            for i in range(64):
                m = oracle_py.leaf_material(bits, padded, i)
                assert (m == mat) if present[i] else (m is None)


def test_leaf_builder_varying_value():
    # gpu_base_dag_test.cpp:62-95.
    present = [1 if i % 3 else 0 for i in range(64)]
    mats = [i % 11 for i in range(64)]
    leaf = oracle_py.leaf_build(4, present, mats)
    assert oracle_py.leaf_size(4, leaf) == len(leaf)
    for i in range(64):
        m = oracle_py.leaf_material(4, leaf, i)
        assert (m == i % 11) if i % 3 else (m is None)


def test_voxel_bit_is_morton_low_bits():
    # A single set voxel lands at leaf bit morton3D(x,y,z) & 63 (tracer.cu:237-241).
    d = oracle_py.OracleDag(6, 4)
    rng = np.random.default_rng(3)
    for x, y, z in rng.integers(0, 64, size=(40, 3)):
        d.root = 0xFFFFFFFF
        d.edit(BOX_ADD, 5, (float(x), float(y), float(z)), 0.0)
        _, _, dump = d.canonical(dump=True)
        leaf_words = dump[3:3 + int(dump[2])]  # level 4 holds the one leaf
        bits = int(leaf_words[0]) | int(leaf_words[1]) << 32
        assert bits == 1 << (morton3d(int(x), int(y), int(z)) & 63)


def _grid(side):
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"), -1)
    return g.reshape(-1, 3).astype(np.uint32)


def _voxels(d, side):
    v = d.query(_grid(side))
    return v.reshape(side, side, side)


def _terrain(levels, seed=11, base=None, amp=None, freq=None, octaves=4):
    side = 1 << levels
    base = side * 0.45 if base is None else base
    amp = side * 0.2 if amp is None else amp
    freq = 4.0 / side if freq is None else freq
    d = oracle_py.OracleDag(levels, 4)
    d.build_terrain(seed, base, amp, freq, octaves, 1)
    h = oracle_py.terrain_heights(seed, base, amp, freq, octaves, 0, 0, side, side)
    return d, h


def _parse(dump):
    levels, p = {}, 0
    while p < len(dump) - 1:
        L, n = int(dump[p]), int(dump[p + 1])
        p += 2
        items = []
        for _ in range(n):
            s = int(dump[p])
            items.append([int(x) for x in dump[p + 1:p + 1 + s]])
            p += 1 + s
        levels[L] = items
    return levels


def _leaf_full(w):
    return w[0] == 0xFFFFFFFF and w[1] == 0xFFFFFFFF and len(set(w[2:])) == 1


def _flag_rule_violations(dump, leaf_level):
    """(level, item) of nodes whose full flag disagrees with the factory rule."""
    lv = _parse(dump)
    bad = []
    for L in range(leaf_level - 1, -1, -1):
        below = lv[L + 1]
        for it in lv[L]:
            hdr, kids = it[0], it[1:]
            if L + 1 == leaf_level:
                full = [_leaf_full(below[k]) for k in kids]
                mats = {below[k][2] & 15 for k in kids}
            else:
                full = [(below[k][0] >> 8) & 1 == 1 for k in kids]
                mats = {(below[k][0] >> 9) & 15 for k in kids}
            rule = (hdr & 0xFF) == 0xFF and all(full) and len(mats) == 1
            if bool((hdr >> 8) & 1) != rule:
                bad.append((L, it))
    return bad


def test_terrain_matches_brute_force_and_factory_rule():
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    vox = _voxels(d, side)
    ys = np.arange(side)[None, :, None]
    expect = ys <= h.T[:, None, :]  # h[z, x] -> [x, 1, z]
    assert np.array_equal(vox != NONE, expect)
    assert np.all(vox[expect] == 1)
    _, _, dump = d.canonical(dump=True)
    assert _flag_rule_violations(dump, levels - 2) == []


def _sphere_overlap_grid(c, r, side):
    # Per-voxel MyGpuSphereEditorBase::test_node_overlap in float32 (my_gpu_dag_editors.h:420-440).
    c = np.float32(c[0]) + np.float32(0.5), np.float32(c[1]) + np.float32(0.5), np.float32(c[2]) + np.float32(0.5)
    rr = np.float32(max(r, 0.1))
    dist = np.full((side, side, side), rr * rr, np.float32)
    coords = np.arange(side, dtype=np.float32)
    for axis in range(3):
        mn = coords
        mx = coords + np.float32(1.0)
        d_lo = np.where(c[axis] < mn, (c[axis] - mn) * (c[axis] - mn), np.float32(0.0)).astype(np.float32)
        d_hi = np.where(c[axis] > mx, (c[axis] - mx) * (c[axis] - mx), np.float32(0.0)).astype(np.float32)
        sub = np.where(c[axis] < mn, d_lo, d_hi).astype(np.float32)
        shape = [1, 1, 1]
        shape[axis] = side
        dist = (dist - sub.reshape(shape)).astype(np.float32)
    return dist > np.float32(0.0)


def _box_overlap_grid(c, r, side):
    lo = [np.float32(x) - np.float32(r) for x in c]
    hi = [np.float32(x) + (np.float32(1.0) if r == 0 else np.float32(r)) for x in c]
    coords = np.arange(side, dtype=np.float32)
    m = np.ones((side, side, side), bool)
    for a in range(3):
        ok = (lo[a] < coords + np.float32(1.0)) & (hi[a] > coords)
        shape = [1, 1, 1]
        shape[a] = side
        m &= ok.reshape(shape)
    return m


def test_edits_match_brute_force_voxelisation():
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    vox = _voxels(d, side).astype(np.int64)
    rng = np.random.default_rng(21)
    for step in range(14):
        kind = [SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, BOX_REMOVE, PAINT][step % 5]
        x, z = (int(v) for v in rng.integers(8, side - 8, 2))
        c = (float(x) + 0.25 * (step % 3), float(h[z, x]) + float(rng.integers(-6, 7)), float(z))
        r = float(rng.integers(2, 16)) + (0.5 if step % 4 == 0 else 0.0)
        mat = int(rng.integers(2, 15))
        if kind in (SPHERE_ADD, SPHERE_REMOVE, PAINT):
            ov = _sphere_overlap_grid(c, r, side)
        else:
            ov = _box_overlap_grid(c, r, side)
        applied, _ = d.edit(kind, mat, c, r)
        assert applied == 1
        if kind in (SPHERE_ADD, BOX_ADD):
            vox[ov] = mat
        elif kind in (SPHERE_REMOVE, BOX_REMOVE):
            vox[ov] = NONE
        else:
            vox[ov & (vox != NONE)] = mat
        got = _voxels(d, side).astype(np.int64)
        assert np.array_equal(got, vox), (step, kind, c, r)


def _copy_rule(vox, src, dst, center, r, rot=0):
    """The copy tool voxel by voxel (my_gpu_dag_editors.h:127-253): every
    voxel whose unit box overlaps the destination box [c - r, c + r) (c =
    dst + (center - src), float32) takes the material of the source voxel
    s = v + trunc(src - dst) when that voxel is set (read from the DAG before
    the edit) and keeps its own value otherwise.  rot = k quarter turns about
    +Y instead map s = q + R^k(v - p), R(x, y, z) = (-z - 1, y, x), about the
    destination pivot p = floor(c + 0.5) and the source pivot
    q = p + trunc(src - dst)."""
    f = np.float32
    side = vox.shape[0]
    c = [f(dst[a]) + (f(center[a]) - f(src[a])) for a in range(3)]
    lo = [c[a] - f(r) for a in range(3)]
    hi = [c[a] + (f(1.0) if r == 0 else f(r)) for a in range(3)]
    off = [int(-f(dst[a]) + f(src[a])) for a in range(3)]
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    inbox = np.ones(len(g), bool)
    for a in range(3):
        inbox &= (lo[a] < g[:, a] + 1) & (hi[a] > g[:, a])
    v = g[inbox].astype(np.int64)
    if rot == 0:
        s = v + np.array(off)
    else:
        p = np.array([int(np.floor(c[a] + f(0.5))) for a in range(3)])
        q = p + np.array(off)
        u = v - p
        x, y, z = u[:, 0], u[:, 1], u[:, 2]
        for _ in range(rot):
            x, z = -z - 1, x
        s = q + np.stack([x, y, z], 1)
    s = np.maximum(s, 0)
    inside = np.all(s < side, axis=1)
    srcval = np.full(len(v), NONE, np.int64)
    srcval[inside] = vox[s[inside, 0], s[inside, 1], s[inside, 2]]
    out = vox.copy()
    keep = srcval == NONE
    out[v[~keep, 0], v[~keep, 1], v[~keep, 2]] = srcval[~keep]
    return out


COPY_CASES = [
    # (src, dst, center, radius, rotation)
    ((40.0, 60.0, 40.0), (84.0, 62.0, 70.0), (40.0, 60.0, 40.0), 11.0, 0),   # integer offset
    ((40.5, 57.25, 40.75), (80.0, 62.5, 70.0), (41.0, 58.0, 39.5), 9.5, 0),  # fractional offset and box
    ((70.0, 50.0, 30.0), (60.0, 55.0, 36.0), (70.0, 50.0, 30.0), 13.0, 0),   # overlapping source and destination
    ((30.0, 58.0, 90.0), (96.0, 60.0, 30.0), (30.0, 58.0, 90.0), 12.0, 1),
    ((30.0, 58.0, 90.0), (96.0, 60.0, 30.0), (30.0, 58.0, 90.0), 12.0, 2),
    ((30.0, 58.0, 90.0), (96.0, 60.0, 30.0), (31.0, 57.0, 88.0), 10.5, 3),
    ((52.25, 60.5, 47.0), (58.0, 61.0, 52.75), (52.0, 60.0, 47.0), 14.0, 1),  # fractional, overlapping, rotated
    ((20.0, 55.0, 20.0), (120.0, 60.0, 118.0), (20.0, 55.0, 20.0), 15.0, 3),  # destination past the volume edge
]


@pytest.mark.parametrize("case", range(len(COPY_CASES)))
def test_copy_matches_brute_force(case):
    src, dst, center, r, rot = COPY_CASES[case]
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    # Varied materials around the source so that copies carry more than one.
    rng = np.random.default_rng(5 + case)
    for _ in range(6):
        c = tuple(float(center[a]) + float(rng.integers(-10, 11)) for a in range(3))
        d.edit([SPHERE_ADD, BOX_ADD, PAINT][int(rng.integers(0, 3))], int(rng.integers(2, 16)), c,
               float(rng.integers(3, 8)))
    before = _voxels(d, side).astype(np.int64)
    want = _copy_rule(before, src, dst, center, r, rot)
    assert not np.array_equal(want, before)
    d.edit(COPY, 0, center, r, src, dst, rot)
    got = _voxels(d, side).astype(np.int64)
    assert np.array_equal(got, want), int((got != want).sum())


@pytest.mark.parametrize("k", [1, 2, 3])
def test_copy_rotations_compose(k):
    # A cube with an integer centre and radius maps onto itself under the
    # quarter turns.  Copied A -> B with k turns and B -> C with 4 - k turns,
    # both into emptied cubes, C holds exactly A's voxels: the two maps
    # compose to a translation.  B is A turned: its voxel u (from the cube's
    # centre) is A's voxel R^k(u).
    levels = 7
    side = 1 << levels
    d, _ = _terrain(levels)
    rng = np.random.default_rng(40 + k)
    for _ in range(5):
        c = tuple(float(v) for v in rng.integers(24, 40, 3))
        d.edit(BOX_ADD, int(rng.integers(2, 16)), (c[0], c[1] + 24.0, c[2]), float(rng.integers(2, 5)))
    r = 10
    A, B, C = (32, 58, 32), (90, 70, 40), (40, 96, 96)
    for cen in (B, C):
        d.edit(BOX_REMOVE, 0, tuple(float(v) for v in cen), float(r))
    d.edit(COPY, 0, tuple(float(v) for v in A), float(r), tuple(float(v) for v in A), tuple(float(v) for v in B), k)
    d.edit(COPY, 0, tuple(float(v) for v in B), float(r), tuple(float(v) for v in B), tuple(float(v) for v in C),
           (4 - k) % 4)
    vox = _voxels(d, side)

    def cube(cen):
        return vox[cen[0] - r:cen[0] + r, cen[1] - r:cen[1] + r, cen[2] - r:cen[2] + r]

    a, b, c = cube(A), cube(B), cube(C)
    assert (a != NONE).any() and (a == NONE).any()
    assert np.array_equal(c, a)
    u = np.arange(-r, r)
    ux, uy, uz = np.meshgrid(u, u, u, indexing="ij")
    x, z = ux, uz
    for _ in range(k):
        x, z = -z - 1, x
    assert np.array_equal(b, a[x + r, uy + r, z + r])


def test_paint_overlap_distance_bounds():
    # gpu_hash_dag_tests.cpp:168-196.
    center = (3599.0, 8440.0, 5045.0)
    radius = 15.0
    for x in range(3599 - 15, 3599 + 16, 3):
        for y in range(8440 - 15, 8440 + 16, 3):
            for z in range(5045 - 15, 5045 + 16, 3):
                mn = (float(x), float(y), float(z))
                mx = (x + 1.0, y + 1.0, z + 1.0)
                dist = np.sqrt((center[0] - x) ** 2 + (center[1] - y) ** 2 + (center[2] - z) ** 2)
                ov = oracle_py.sphere_overlap(center, radius, mn, mx)
                if dist < radius - 1:
                    assert ov
                elif dist > radius + 1:
                    assert not ov


GATE_CASE = dict(levels=6, terrain=(7, 30.0, 0.0, 1 / 64, 1, 1),
                 edits=[(BOX_ADD, 1, (28.0, 39.0, 37.0), 15.0), (SPHERE_ADD, 1, (20.0, 46.0, 12.0), 6.0),
                        (BOX_ADD, 1, (12.0, 63.0, 55.0), 3.0)])


def test_fully_filled_gate_leaves_a_flagless_full_node():
    # my_gpu_hash_dag_edits.cu:337: discovery at a level runs only if the
    # level below found a full item, so a node whose 8 children are all full
    # can stay unflagged.  Factory rule on the same voxels flags it.
    c = GATE_CASE
    d = oracle_py.OracleDag(c["levels"], 4)
    d.build_terrain(*c["terrain"])
    for e in c["edits"]:
        assert d.edit(*e)[0] == 1
    side = 1 << c["levels"]
    vox = _voxels(d, side)
    f = oracle_py.OracleDag(c["levels"], 4)
    oracle_py.build_dense(f, np.where(vox == NONE, 255, vox).astype(np.uint8))
    assert np.array_equal(_voxels(f, side), vox)
    bad = _flag_rule_violations(d.canonical(dump=True)[2], c["levels"] - 2)
    assert len(bad) == 1 and bad[0][0] == 2 and (bad[0][1][0] & 0xFF) == 0xFF
    assert _flag_rule_violations(f.canonical(dump=True)[2], c["levels"] - 2) == []
    assert d.canonical()[1][2] == f.canonical()[1][2] + 1


def test_edit_without_leaves_changes_nothing():
    # my_gpu_hash_dag_edits.cu:651: a box exactly covering an octant makes no leaf.
    d, _ = _terrain(6)
    before = d.canonical()[0]
    applied, counts = d.edit(BOX_ADD, 7, (16.0, 16.0, 16.0), 16.0)
    assert applied == 0 and counts[4] == 0
    assert d.canonical()[0] == before


def _march(vox, o, dvec, side):
    """First solid voxel along the ray (Amanatides-Woo in double) and its min
    distance to a voxel edge at the entry point."""
    o = np.array(o, float)
    dvec = np.array(dvec, float)
    tmin, tmax = 0.0, np.inf
    for a in range(3):
        if dvec[a] == 0:
            if not 0 <= o[a] <= side:
                return None, 1.0
            continue
        t1, t2 = (0 - o[a]) / dvec[a], (side - o[a]) / dvec[a]
        tmin, tmax = max(tmin, min(t1, t2)), min(tmax, max(t1, t2))
    if tmin >= tmax:
        return None, 1.0
    p = o + (tmin + 1e-9) * dvec
    cell = np.clip(np.floor(p).astype(int), 0, side - 1)
    step = np.sign(dvec).astype(int)
    nxt = np.where(step > 0, cell + 1, cell).astype(float)
    with np.errstate(divide="ignore"):
        tmax_a = np.where(dvec != 0, (nxt - o) / dvec, np.inf)
        tdelta = np.where(dvec != 0, np.abs(1.0 / dvec), np.inf)
    t = tmin
    while np.all((cell >= 0) & (cell < side)):
        if vox[cell[0], cell[1], cell[2]]:
            q = o + t * dvec
            frac = np.abs(q - np.round(q))
            return tuple(int(v) for v in cell), float(np.sort(frac)[1])
        a = int(np.argmin(tmax_a))
        t = tmax_a[a]
        cell[a] += step[a]
        tmax_a[a] += tdelta[a]
    return None, 1.0


def test_tracer_hits_first_set_voxel():
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    d.edit(SPHERE_REMOVE, 0, (60.0, float(h[60, 60]), 60.0), 12.0)
    vox = _voxels(d, side) != NONE
    pos = (-0.2 * side, 0.9 * side, -0.15 * side)
    tgt = np.array((0.55 * side, 0.3 * side, 0.5 * side))
    f = tgt - np.array(pos)
    f /= np.linalg.norm(f)
    r = np.cross(f, (0.0, 1.0, 0.0))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    rot = np.concatenate([-r, u, f])
    W = H = 40
    hits = d.trace(pos, rot, 60.0, (0, 0, 0), (side, side, side), W, H)
    fov = 60.0 / 2 * np.pi / 180
    right, up, fwd = -rot[0:3], rot[3:6], rot[6:9]
    X, Y, Z = right * np.sin(fov) * (W / H), up * np.sin(fov), fwd * np.cos(fov)
    bl, br, tl = np.array(pos) + Z - Y - X, np.array(pos) + Z - Y + X, np.array(pos) + Z + Y - X
    agree = ambiguous = 0
    for py in range(H):
        for px in range(W):
            hit = hits[(H - 1 - py) * W + px]
            if hit["material"] != NONE:
                assert vox[hit["x"], hit["y"], hit["z"]]
            dvec = bl + px * (br - bl) / W + py * (tl - bl) / H - np.array(pos)
            ref, edge = _march(vox, pos, dvec / np.linalg.norm(dvec), side)
            got = None if hit["material"] == NONE else (int(hit["x"]), int(hit["y"]), int(hit["z"]))
            if got == ref:
                agree += 1
            else:
                assert edge < 1e-3, (px, py, got, ref)
                ambiguous += 1
    assert agree >= 0.98 * W * H


def _fma64(a, b, c):
    """Correctly rounded double fma (exact rational, one rounding)."""
    from fractions import Fraction
    return float(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def _primary_rays(pos, rot, fov_deg, bmin, bmax, levels, W, H):
    """get_trace_params (dag_tracer.cu:59-101) in double, then the per-pixel
    ray of trace_paths (tracer.cu:402-405): (rayMin + px*ddx + py*ddy) - cam
    with each add fused (nvcc --fmad=true), normalised as (1/length(a)) * a
    and rounded to float."""
    right, up, fwd = -np.array(rot[0:3]), np.array(rot[3:6]), np.array(rot[6:9])
    fov = fov_deg / 2.0 * (np.pi / 180.0)
    aspect = W / H
    cam, rmin, ddx, ddy = [], [], [], []
    for k in range(3):
        X, Y, Z = right[k] * np.sin(fov) * aspect, up[k] * np.sin(fov), fwd[k] * np.cos(fov)
        bl, br, tl = pos[k] + Z - Y - X, pos[k] + Z - Y + X, pos[k] + Z + Y - X
        tr = -bmin[k]
        sc = float(1 << levels) / (bmax[k] - bmin[k])
        cam.append((pos[k] + tr) * sc)
        rmin.append((bl + tr) * sc)
        ddx.append(((br + tr) * sc - rmin[k]) * (1.0 / W))
        ddy.append(((tl + tr) * sc - rmin[k]) * (1.0 / H))
    rays = {}
    for py in range(H):
        for px in range(W):
            a = [_fma64(py, ddy[k], _fma64(px, ddx[k], rmin[k])) - cam[k] for k in range(3)]
            inv = 1.0 / np.sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2])
            d = [np.float32(inv * a[k]) for k in range(3)]
            rays[(px, py)] = ([np.float32(c) for c in cam], d)
    return rays


def test_tracer_hit_t_and_position_are_the_entry_face():
    # The hit record's t is the ray's entry distance into the hit voxel's box
    # (the largest of the per-axis near-slab distances), and its position is
    # origin + t * dir on that entry face.  Both are checked against exact
    # rational arithmetic over the float ray, with the rounding of t's three
    # float operations (1/d, f - o, product) as the tolerance.
    from fractions import Fraction
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    d.edit(SPHERE_REMOVE, 0, (60.0, float(h[60, 60]), 60.0), 12.0)
    pos = (-0.2 * side, 0.9 * side, -0.15 * side)
    tgt = np.array((0.55 * side, 0.3 * side, 0.5 * side))
    f = tgt - np.array(pos)
    f /= np.linalg.norm(f)
    r = np.cross(f, (0.0, 1.0, 0.0))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    rot = np.concatenate([-r, u, f])
    W = H = 24
    hits = d.trace(pos, rot, 60.0, (0, 0, 0), (side, side, side), W, H)
    rays = _primary_rays(pos, rot, 60.0, (0, 0, 0), (side, side, side), levels, W, H)
    eps = Fraction(1, 1 << 24)
    checked = 0
    for py in range(H):
        for px in range(W):
            hit = hits[(H - 1 - py) * W + px]
            if hit["material"] == NONE:
                continue
            o, dv = rays[(px, py)]
            box = (int(hit["x"]), int(hit["y"]), int(hit["z"]))
            near = []
            for a in range(3):
                assert dv[a] != 0
                lo, hi = Fraction(box[a]) - Fraction(float(o[a])), Fraction(box[a] + 1) - Fraction(float(o[a]))
                t1, t2 = lo / Fraction(float(dv[a])), hi / Fraction(float(dv[a]))
                near.append(min(t1, t2))
            t_exact = max(near)
            axis = near.index(t_exact)
            t = Fraction(float(hit["t"]))
            assert abs(t - t_exact) <= 4 * eps * abs(t_exact), (px, py, float(t), float(t_exact))
            p = [Fraction(float(hit[k])) for k in ("px", "py", "pz")]
            for a in range(3):
                scale = abs(t_exact * Fraction(float(dv[a]))) + abs(Fraction(float(o[a]))) + 1
                tol = 8 * eps * scale
                if a == axis:
                    face = box[a] if dv[a] > 0 else box[a] + 1
                    assert abs(p[a] - face) <= tol, (px, py, a, float(p[a]), face)
                else:
                    assert box[a] - tol <= p[a] <= box[a] + 1 + tol, (px, py, a, float(p[a]), box)
            checked += 1
    assert checked > W * H // 3


def test_sun_direction_is_the_float_normalisation():
    # sun_direction() = normalize(float3(0.3f, 1, 0.5f)) (tracer.cu:600-603)
    # with normalize(a) = (1 / length(a)) * a in float (cuda_math.h:115-118):
    # a unit vector within float rounding, parallel to (0.3, 1, 0.5), and not
    # the double-normalised vector rounded once in at least one component.
    L = oracle_py.sun_dir()
    v = np.array([np.float32(0.3), 1.0, 0.5])
    assert abs(float(np.dot(L.astype(np.float64), L.astype(np.float64))) - 1.0) < 4e-7
    ref = v / np.linalg.norm(v)
    assert np.all(np.abs(L.astype(np.float64) - ref) <= 2 * np.spacing(np.float32(1.0)))
    f = np.float32
    dot = (f(0.3) * f(0.3) + f(1.0) * f(1.0)) + f(0.5) * f(0.5)
    inv = f(1.0) / np.sqrt(dot, dtype=np.float32)
    assert np.array_equal(L, np.array([inv * f(0.3), inv * f(1.0), inv * f(0.5)], np.float32))


def test_banded_terrain_voxels_follow_heights_and_bands():
    # The factory-rule DAG over height bands holds exactly the voxel field:
    # set iff y <= h(x, z), material by band and column tile.
    L, bits, band, seed = 8, 8, 4, 0xBA5E
    side = 1 << L
    base, amp, freq, octaves, mat = 128.0, 50.0, 1 / 64, 5, 7
    d = oracle_py.OracleDag(L, bits)
    d.build_terrain(0x5EED, base, amp, freq, octaves, mat, band, seed)
    h = oracle_py.terrain_heights(0x5EED, base, amp, freq, octaves, 0, 0, side, side)
    rng = np.random.default_rng(11)
    xz = rng.integers(0, side, size=(400, 2))
    pts = np.array([(x, y, z) for x, z in xz for y in range(side)], np.uint32)
    got = d.query(pts)
    x, y, z = (pts[:, i].astype(np.uint64) for i in range(3))
    t = (np.uint64(seed) ^ (((x >> band) * 0x9E3779B1) & 0xFFFFFFFF) ^ (((z >> band) * 0x85EBCA77) & 0xFFFFFFFF))
    t = ((t ^ (t >> 15)) * 0x2C1B3C6D) & 0xFFFFFFFF
    t ^= t >> 12
    want_mat = (mat + (y >> band) + (t & 3)) & ((1 << bits) - 1)
    solid = y.astype(np.int64) <= h[z.astype(np.int64), x.astype(np.int64)]
    want = np.where(solid, want_mat, 0xFFFFFFFF).astype(np.uint32)
    assert solid.any() and (~solid).any()
    assert len(np.unique(want_mat[solid])) > 4
    assert np.array_equal(got, want)


@pytest.mark.parametrize("case", range(len(COPY_CASES)))
def test_sampled_copy_check_agrees_with_the_oracle(case):
    # The large-scene leg checks copies by sampling: destination voxels from
    # workloads.copy_box_voxels, their sources from copy_source_voxels, and
    # the rule new = source if set else old.  That check must accept the
    # oracle's copy and reject a copy that ignores the rotation.
    src, dst, center, r, rot = COPY_CASES[case]
    levels = 7
    side = 1 << levels
    d, _ = _terrain(levels)
    rng = np.random.default_rng(5 + case)
    for _ in range(6):
        c = tuple(float(center[a]) + float(rng.integers(-10, 11)) for a in range(3))
        d.edit([SPHERE_ADD, BOX_ADD, PAINT][int(rng.integers(0, 3))], int(rng.integers(2, 16)), c,
               float(rng.integers(3, 8)))
    c = [np.float32(dst[a]) + (np.float32(center[a]) - np.float32(src[a])) for a in range(3)]
    v = workloads.copy_box_voxels([float(x) for x in c], r, 3000, np.random.default_rng(case), levels)
    sv = workloads.copy_source_voxels(v, center, r, src, dst, rot)
    inside = np.all(sv < side, axis=1)
    v, sv = v[inside], sv[inside]
    old = d.query(v.astype(np.uint32))
    srcv = d.query(sv.astype(np.uint32))
    d.edit(COPY, 0, center, r, src, dst, rot)
    new = d.query(v.astype(np.uint32))
    assert np.array_equal(new, np.where(srcv != 0xFFFFFFFF, srcv, old))
    if rot:
        wrong = workloads.copy_source_voxels(v, center, r, src, dst, 0)
        srcw = d.query(np.minimum(wrong, side - 1).astype(np.uint32))
        assert not np.array_equal(new, np.where(srcw != 0xFFFFFFFF, srcw, old))


def test_band_materials_match_the_oracle_terrain():
    L, bits, band, seed = 8, 8, 4, 0xBA5E
    side = 1 << L
    base, amp, freq, octaves, mat = 128.0, 50.0, 1 / 64, 5, 7
    d = oracle_py.OracleDag(L, bits)
    d.build_terrain(0x5EED, base, amp, freq, octaves, mat, band, seed)
    h = oracle_py.terrain_heights(0x5EED, base, amp, freq, octaves, 0, 0, side, side)
    rng = np.random.default_rng(12)
    x, z = rng.integers(0, side, 5000), rng.integers(0, side, 5000)
    y = np.minimum(h[z, x], side - 1) - rng.integers(0, 40, 5000)
    keep = y >= 0
    x, y, z = x[keep], y[keep], z[keep]
    got = d.query(np.stack([x, y, z], 1).astype(np.uint32))
    assert np.array_equal(got, workloads.band_materials(x, y, z, mat, band, seed, bits))
