"""The oracle's colour, blur and lighting passes (oracle/oracle_shading.cpp),
pinned on the CPU by what the mathematics fixes rather than by retyping:

  - the sRGB transfer functions against the textbook piecewise curves in
    double precision (IEC 61966-2-1), and their 8-bit round trip;
  - float3_to_rgb888's truncation and clamping at exact values;
  - the blur weights against the Gaussian the reference writes (mean 4.5,
    the sigma passed as the variance, normalised), their sum, the blur of a
    constant image (the weight sum inside, partial sums at the borders) and
    separability (a blurred impulse is the outer product of the weights);
  - lighting: a miss is the sky made sRGB; with occlusion 1 and no sun a hit
    is 0.5 * sky * linear(colour) made sRGB; occlusion scales it; fog 0
    changes nothing and heavy fog far away tends to the fog colour;
  - path tracing: an empty scene sums the sky once per sample and a zero
    environment gives black; one bounce on a flat floor is the closed-form
    sun term N.L / pi * Kd * sunColor; sums accumulate across calls.
"""
import ctypes

import numpy as np

import oracle_py


def _srgb_to_linear(c):
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


def _linear_to_srgb(c):
    return c * 12.92 if c <= 0.0031308 else 1.055 * abs(c) ** (1 / 2.4) - 0.055


def test_srgb_transfer_functions_follow_the_standard_curves():
    for v in range(256):
        c = np.float32(v / 255.0)
        lin = oracle_py.srgb_to_linear(c)
        assert abs(lin - _srgb_to_linear(float(c))) <= 2e-7 * max(1.0, lin) + 1e-9
    for c in np.linspace(0.0, 1.0, 1001, dtype=np.float32):
        s = oracle_py.linear_to_srgb(c)
        assert abs(s - _linear_to_srgb(float(c))) <= 3e-7
    # Textbook points: linear 0.5 -> sRGB 0.7353569; sRGB 0.5 -> linear 0.2140411.
    assert abs(oracle_py.linear_to_srgb(0.5) - 0.7353569) < 1e-6
    assert abs(oracle_py.srgb_to_linear(0.5) - 0.2140411) < 1e-6
    # 8-bit values survive linear and back within one count (truncation).
    for v in range(256):
        back = oracle_py.rgb888([oracle_py.linear_to_srgb(oracle_py.srgb_to_linear(np.float32(v / 255.0)))] * 3)
        assert v - 1 <= (back & 0xFF) <= v


def test_rgb888_truncates_and_clamps():
    assert oracle_py.rgb888([1.0, 0.5, 0.0]) == 0xFF00_7FFF  # 255, 127 (127.5 truncated), 0, alpha 255
    assert oracle_py.rgb888([1.5, -0.2, 2.0 / 255.0]) == 0xFF02_00FF
    assert oracle_py.rgb888([0.0, 0.0, 0.0]) == 0xFF00_0000


def test_blur_weights_are_the_written_gaussian():
    w = oracle_py.blur_weights().astype(np.float64)
    x = np.arange(9, dtype=np.float64)
    sigma = 0.3 * 3 + 0.8
    g = np.exp(-0.5 * (x - 4.5) ** 2 / sigma)  # normalisation cancels
    assert np.allclose(w, g / g.sum(), rtol=2e-6, atol=0)
    assert abs(w.sum() - 1.0) < 1e-6
    assert w[4] == w[5] and w[3] == w[6] and w[0] != w[8]  # centred on 4.5: the reference's off-centre kernel


def test_blur_of_constant_and_impulse_images():
    W, H = 24, 16
    w = oracle_py.blur_weights().astype(np.float64)
    out = oracle_py.blur(np.full(W * H, 0.75, np.float32), W, H).reshape(H, W).astype(np.float64)
    inner = out[4:-4, 4:-4]
    assert np.allclose(inner, 0.75 * w.sum() ** 2, rtol=1e-6)
    # The first column sees the horizontal taps 4..8 only.
    assert np.allclose(out[8, 0], 0.75 * w[4:].sum() * w.sum(), rtol=1e-6)
    img = np.zeros((H, W), np.float32)
    img[8, 12] = 1.0
    b = oracle_py.blur(img.reshape(-1), W, H).reshape(H, W).astype(np.float64)
    # out(y, x) = w[8 - y + ...]: the impulse spreads as the reversed weights.
    want = np.zeros((H, W))
    for dy in range(-4, 5):
        for dx in range(-4, 5):
            want[8 - dy, 12 - dx] = w[dy + 4] * w[dx + 4]
    assert np.allclose(b, want, rtol=1e-6, atol=1e-12)


def _hits(recs):
    out = np.zeros(len(recs), oracle_py.HIT_DTYPE)
    for i, (x, y, z, m, p) in enumerate(recs):
        out[i] = (x, y, z, m, 1.0, p[0], p[1], p[2])
    return out


def test_colors_checker_and_palette():
    pal = np.array([[0.2, 0.4, 0.6], [1.0, 0.0, 0.5]], np.float32)
    hits = _hits([(0, 0, 0, 0xFFFFFFFF, (0, 0, 0)), (3, 5, 9, 1, (0, 0, 0)), (3, 5, 8, 7, (0, 0, 0)),
                  (3, 5, 9, 7, (0, 0, 0))])
    c = oracle_py.colors(hits, pal)
    assert c[0] == 0xFF000000  # no surface: black
    assert c[1] == oracle_py.rgb888(pal[1])
    assert c[2] == oracle_py.rgb888([1.0, 0.0, 1.0])  # 3^5^8 even: (1, 0, 1)
    assert c[3] == oracle_py.rgb888([1.0, 1.0, 0.0])  # 3^5^9 odd: (1, 1, 0)


def test_lighting_of_sky_and_ambient_only_surfaces():
    sky = np.array([187, 242, 250], np.float64) / 255.0
    hits = _hits([(0, 0, 0, 0xFFFFFFFF, (0, 0, 0)), (1, 2, 3, 0, (1.5, 2.0, 3.5)), (1, 2, 3, 0, (1.5, 2.0, 3.5)),
                  (1, 2, 3, 0, (1.5, 2.0, 3.5))])
    col = np.array([0, 0xFF3080C0, 0xFF3080C0, 0xFF3080C0], np.uint32)
    ao = np.array([0.0, 1.0, 0.25, 1.0], np.float32)
    sun = np.array([0.0, 0.7, 0.7, 0.7], np.float32)
    cam = ((-5.0, 30.0, -5.0), (-1, 0, 0, 0, 1, 0, 0, 0, 1), 60.0, (0, 0, 0), (64, 64, 64), 6)
    out = oracle_py.lighting(hits, 2, 2, col, ao, sun, *cam, sun_brightness=0.0)
    out_sun = oracle_py.lighting(hits, 2, 2, col, ao, sun, *cam, sun_brightness=2.0)
    chan = lambda v, k: (int(v) >> (8 * k)) & 0xFF
    for k in range(3):
        assert abs(chan(out[0], k) - int(_linear_to_srgb(sky[k]) * 255)) <= 1
        kd = _srgb_to_linear(((0xFF3080C0 >> (8 * k)) & 0xFF) / 255.0)
        assert abs(chan(out[1], k) - int(_linear_to_srgb(0.5 * sky[k] * kd) * 255)) <= 1
        assert abs(chan(out[2], k) - int(_linear_to_srgb(0.25 * 0.5 * sky[k] * kd) * 255)) <= 1
        assert chan(out_sun[1], k) >= chan(out[1], k)  # the sun only adds light
    assert out_sun[0] == out[0]  # the sky has no sun term
    fog0 = oracle_py.lighting(hits, 2, 2, col, ao, sun, *cam, fog_density=0.0)
    assert np.array_equal(fog0, out)
    far = _hits([(1, 2, 3, 0, (1e6, 1e6, 1e6))] * 4)
    fogged = oracle_py.lighting(far, 2, 2, col, None, None, *cam, fog_density=1e4)
    for v in fogged:
        for k in range(3):
            assert chan(v, k) >= int(_linear_to_srgb(sky[k]) * 255) - 1  # the fog colour lies between sky and white


def _floor_dag(levels, top, material):
    # A flat floor: every voxel with y <= top set, one material.  The box
    # editor covers [c - r, c + r] on every axis; its y range ends inside
    # voxel `top` (an edge on a leaf boundary would make no partial leaf, and
    # the reference skips edits without one).
    o = oracle_py.OracleDag(levels, 4)
    half = float(1 << levels) / 2.0
    o.edit(2, material, (half, top + 0.5 - half, half), half)
    return o


def test_path_tracing_of_an_empty_scene_is_the_sky():
    levels, W, H = 6, 8, 8
    o = oracle_py.OracleDag(levels, 4)
    cam = ((32.0, 40.0, -20.0), (-1, 0, 0, 0, 1, 0, 0, 0, 1), 60.0, (0, 0, 0), (64, 64, 64))
    img, acc = o.path_trace(*cam, W, H, None, np.zeros((1, 3)), env=1.0, samples=3, depth=4, integrate=True)
    sky = np.array([187, 242, 250], np.float32) / np.float32(255.0)
    assert np.allclose(acc.reshape(-1, 3), 3 * sky, rtol=1e-6)
    want = oracle_py.rgb888([oracle_py.linear_to_srgb(v) for v in sky])
    assert (img == want).all()
    dark, _ = o.path_trace(*cam, W, H, None, np.zeros((1, 3)), env=0.0, samples=2, depth=3, integrate=True)
    assert (dark == 0xFF000000).all()


def test_path_tracing_one_bounce_is_sun_light_on_the_floor():
    # Depth 1 on a flat floor seen from above: a lit floor pixel gets
    # max(N.L, 0) / pi * Kd * sunColor * env (next-event estimation), with
    # N = +y and L the sun direction; the continuation adds nothing more.
    levels, W, H = 6, 16, 16
    o = _floor_dag(levels, 9, 3)
    L = np.zeros(3, np.float32)
    oracle_py.dlib().ora_sun_dir(L.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    rot = (-1, 0, 0, 0, 0, 1, 0, -1, 0)  # looking down -y
    cam = ((32.0, 60.0, 32.0), rot, 30.0, (0, 0, 0), (64, 64, 64))
    pal = np.zeros((16, 3), np.float32)
    pal[3] = (0.8, 0.5, 0.2)
    hits = o.trace(*cam, W, H)
    assert (hits["y"] == 9).all()  # every ray hits the floor's top voxels
    colors = oracle_py.colors(hits, pal)
    img, acc = o.path_trace(*cam, W, H, colors, pal, env=1.0, samples=1, depth=1, integrate=False)
    kd = [_srgb_to_linear(((int(colors[0]) >> (8 * k)) & 0xFF) / 255.0) for k in range(3)]
    sun = (0.95, 0.94, 0.47)
    want = np.array([float(L[1]) / np.pi * kd[k] * sun[k] for k in range(3)])
    a = acc.reshape(-1, 3)
    lit = np.all(np.isclose(a, want, rtol=2e-6, atol=0), axis=1)
    # The camera's centre row has a direction component of exactly 0, where
    # createSurfaceInteraction's entry axis is a side face, away from the sun.
    assert lit.mean() > 0.9 and np.all(a[~lit] == 0)


def test_path_tracing_accumulates_across_calls():
    levels, W, H = 6, 12, 8
    o = _floor_dag(levels, 21, 5)
    rot = (-1, 0, 0, 0, 1, 0, 0, 0, 1)
    cam = ((32.0, 30.0, -10.0), rot, 60.0, (0, 0, 0), (64, 64, 64))
    pal = np.full((16, 3), 0.6, np.float32)
    img1, acc1 = o.path_trace(*cam, W, H, None, pal, env=1.0, samples=2, depth=3, integrate=True, seed=5)
    img2, acc2 = o.path_trace(*cam, W, H, None, pal, env=1.0, samples=3, depth=3, integrate=True, seed=6, accum=acc1,
                              accumulated=2)
    _, solo = o.path_trace(*cam, W, H, None, pal, env=1.0, samples=3, depth=3, integrate=True, seed=6)
    assert np.allclose(acc2, acc1 + solo, rtol=1e-6, atol=1e-7)
    assert (acc2 >= acc1).all()
