"""Pins for the oracle's restatement of the shadow and ambient-occlusion
passes (tracer.cu:274-351, 600-691; utils.h:45-82)."""
import numpy as np

import oracle_py
from test_oracle_dag import SPHERE_REMOVE, _march, _terrain, _voxels

M64 = (1 << 64) - 1
MULT = 6364136223846793005


def _lcg(state, inc):
    return (state * MULT + (inc | 1)) & M64


def test_pcg32_matches_published_demo_vector():
    # pcg32_srandom_r(&rng, 42, 54) from PCG's reference distribution:
    # state = 0, step, state += 42, step; then six outputs.
    inc = (54 << 1) | 1
    st = _lcg(_lcg(0, inc) + 42, inc)
    got = oracle_py.pcg32_raw(st, inc, 6)
    assert [int(x) for x in got] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_reference_seeding_drops_two_outputs():
    # PCG_RNG(stream): inc = stream, state = 0x853c49e6748fea9b, two outputs dropped.
    for stream in (1, 2, 12345, (1 << 40) + 7):
        st = _lcg(_lcg(0x853C49E6748FEA9B, stream), stream)
        assert np.array_equal(oracle_py.pcg32_seeded(stream, 8), oracle_py.pcg32_raw(st, stream, 8))


def test_sincos_turns_against_libm():
    u = np.linspace(0.0, 1.0, 20001)
    s, c = np.array([oracle_py.sincos_turns(float(np.float32(x))) for x in u]).T
    uf = u.astype(np.float32).astype(np.float64)
    assert np.max(np.abs(s - np.sin(2 * np.pi * uf))) < 3e-7
    assert np.max(np.abs(c - np.cos(2 * np.pi * uf))) < 3e-7
    assert oracle_py.sincos_turns(0.0) == (0.0, 1.0)
    assert oracle_py.sincos_turns(0.25) == (1.0, 0.0)


def test_sun_direction_is_the_normalised_reference_vector():
    L = oracle_py.sun_dir().astype(np.float64)
    v = np.array([np.float32(0.3), 1.0, 0.5])
    assert np.allclose(L, v / np.linalg.norm(v), rtol=0, atol=1e-7)


def test_any_hit_walk_against_brute_force_march():
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    d.edit(SPHERE_REMOVE, 0, (60.0, float(h[60, 60]), 60.0), 14.0)
    vox = _voxels(d, side) != 0xFFFFFFFF
    rng = np.random.default_rng(7)
    agree = checked = 0
    for _ in range(1500):
        o = rng.uniform(0.5, side - 0.5, 3)
        if vox[tuple(np.floor(o).astype(int))]:
            continue
        dvec = rng.normal(size=3)
        dvec /= np.linalg.norm(dvec)
        o32, d32 = o.astype(np.float32), dvec.astype(np.float32)
        ref, edge = _march(vox, o32.astype(float), d32.astype(float), side)
        got = d.occluded(o32, d32)
        checked += 1
        if got == (ref is not None):
            agree += 1
        else:
            assert edge < 1e-3, (o, dvec, got, ref)
    assert checked > 500 and agree >= 0.99 * checked


def test_ambient_occlusion_open_floor_and_closed_cavity():
    levels = 7
    side = 1 << levels
    vox = np.full((side, side, side), 0xFF, np.uint8)
    vox[:, :32, :] = 1                      # solid slab below y = 32
    d = oracle_py.OracleDag(levels, 4)
    oracle_py.build_dense(d, vox)
    up = 1 + 2 * 1 + 1                      # surface code: axis y, normal +y
    # Nothing above the slab: every sample escapes.
    assert d.ao_at((64.3, 32.0, 64.7), up, 64, 1.0, 16.0, 99) == 1.0
    # A closed cavity: every sample hits a wall when rays are long enough,
    # none when they are shorter than the distance to any wall.
    vox2 = np.full((side, side, side), 1, np.uint8)
    vox2[40:60, 40:60, 40:60] = 0xFF
    d2 = oracle_py.OracleDag(levels, 4)
    oracle_py.build_dense(d2, vox2)
    p = (50.3, 40.0, 50.7)
    assert d2.ao_at(p, up, 64, 1.0, 64.0, 5) == 0.0
    assert d2.ao_at(p, up, 64, 1.0, 0.5, 5) == 1.0


def test_shadow_values_are_lit_fractions_of_the_sun_direction():
    levels = 7
    side = 1 << levels
    d, h = _terrain(levels)
    pos = (-0.2 * side, 0.9 * side, -0.15 * side)
    tgt = np.array((0.55 * side, 0.3 * side, 0.5 * side))
    f = tgt - np.array(pos)
    f /= np.linalg.norm(f)
    r = np.cross(f, (0.0, 1.0, 0.0))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    rot = np.concatenate([-r, u, f])
    sun, ao = d.shade(pos, rot, 60.0, (0, 0, 0), (side, side, side), 48, 48, 1.0, 4, 16.0, 3)
    L = oracle_py.sun_dir()
    allowed = {0.0} | {float(x) for x in L if x > 0}
    assert set(np.unique(sun).tolist()) <= allowed
    assert set(np.unique(ao).tolist()) <= {0.0, 0.25, 0.5, 0.75, 1.0}
    hits = d.trace(pos, rot, 60.0, (0, 0, 0), (side, side, side), 48, 48)
    miss = (hits["x"] | hits["y"] | hits["z"]) == 0
    assert (sun[miss] == 0).all() and (ao[miss] == 0).all()
    # Part of the terrain is lit, part is in shadow or faces away.
    lit = (sun[~miss] > 0).mean()
    assert 0.05 < lit < 0.95, lit
