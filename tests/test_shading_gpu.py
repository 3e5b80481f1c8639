"""The passes after the primary trace on the GPU against the oracle:
shadows and ambient occlusion (Tracer::trace_shadows / trace_ambient_occlusion,
tracer.cu:613-691), the occlusion blur (BlurKernel, blur.cu:8-85), colours
(trace_colors, tracer.cu:413-596) and lighting (trace_lighting, tracer.cu:693-750)."""
import numpy as np
import pytest

import oracle_py
import svdag_py
import workloads
from test_dag_gpu import L10, _pair


@pytest.mark.gpu
@pytest.mark.parametrize("samples,seed", [(3, 1), (8, 12345)])
def test_shadows_and_ambient_occlusion_match_oracle(samples, seed):
    g, o = _pair()
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    args = (cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 256, 256)
    g.trace(*args)
    sun, _ = g.trace_shadows(1.0)
    ao, _ = g.trace_ambient_occlusion(samples, 1.0, 32.0, seed)
    osun, oao = o.shade(*args, bias=1.0, ao_samples=samples, ao_length=32.0, seed=seed)
    assert np.array_equal(sun, osun)
    assert np.array_equal(ao, oao)
    assert (sun > 0).mean() > 0.05 and (sun == 0).mean() > 0.05
    assert (ao > 0).any() and ((ao > 0) & (ao < 1)).any()
    g.close()


@pytest.mark.gpu
def test_shading_after_edits_uses_the_traced_root():
    # The passes shade the surface of the last trace, with the DAG it traced.
    g, o = _pair()
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    args = (cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 128, 96)
    g.trace(*args)
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
    sun, _ = g.trace_shadows(0.5)
    ao, _ = g.trace_ambient_occlusion(4, 0.5, 16.0, 9)
    osun, oao = o.shade(*args, bias=0.5, ao_samples=4, ao_length=16.0, seed=9)
    assert np.array_equal(sun, osun) and np.array_equal(ao, oao)
    g.close()


@pytest.mark.gpu
def test_shading_needs_a_primary_trace():
    g = svdag_py.MyGPUHashDAG(L10, 4)
    g._trace_pixels = 16
    with pytest.raises(svdag_py.SvdagError):
        g.trace_shadows()
    with pytest.raises(svdag_py.SvdagError):
        g.trace_ambient_occlusion(0)
    g.close()


@pytest.mark.gpu
def test_blur_colors_and_lighting_match_oracle():
    # The frame after the primary trace, as Engine runs it: occlusion, its
    # blur, shadows, colours, lighting.  Blur and colours are bit-exact; the
    # lighting pass uses powf / expf, whose last bit may differ between the
    # GPU's and the host's libraries, so its 8-bit channels may differ by
    # one count (at most 1% of the pixels here), never by more.
    g, o = _pair()
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    g.edit(svdag_py.PAINT, 9, (560.0, 510.0, 560.0), 60.0)
    o.edit(svdag_py.PAINT, 9, (560.0, 510.0, 560.0), 60.0)
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    W, H = 256, 192
    args = (cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], W, H)
    hits, _ = g.trace(*args)
    ohits = o.trace(*args)
    assert np.array_equal(hits.view(np.uint32), ohits.view(np.uint32))
    sun, _ = g.trace_shadows(1.0)
    ao, _ = g.trace_ambient_occlusion(4, 1.0, 32.0, 123)
    osun, oao = o.shade(*args, bias=1.0, ao_samples=4, ao_length=32.0, seed=123)
    blurred, _ = g.blur_ambient_occlusion()
    assert np.array_equal(blurred, oracle_py.blur(oao, W, H))
    # Materials 0..7 have colours; the paint's 9 (and anything above 7) gets the checker.
    pal = np.random.default_rng(4).uniform(0, 1, (8, 3)).astype(np.float32)
    col, _ = g.trace_colors(pal)
    ocol = oracle_py.colors(ohits, pal)
    assert np.array_equal(col, ocol)
    assert (hits["material"] == 9).any() and (hits["material"] < 8).any()
    for apply_ao, apply_sun, fog, bright in [(True, True, 0.0, 0.0), (True, True, 0.0, 1.5), (False, True, 40.0, 1.0),
                                             (True, False, 0.0, 0.0)]:
        lit, _ = g.trace_lighting(apply_ao, apply_sun, fog, bright)
        olit = oracle_py.lighting(ohits, W, H, ocol, oracle_py.blur(oao, W, H) if apply_ao else None,
                                  osun if apply_sun else None, cam["position"], rot, cam["fov"], cam["bmin"],
                                  cam["bmax"], L10, sun_brightness=bright, fog_density=fog)
        d = np.abs(lit.view(np.uint8).astype(np.int16) - olit.view(np.uint8).astype(np.int16))
        assert d.max() <= 1, (apply_ao, apply_sun, fog, bright)
        assert (d.reshape(-1, 4).max(axis=1) > 0).mean() <= 0.01, (apply_ao, apply_sun, fog, bright)
    g.close()


@pytest.mark.gpu
def test_lighting_needs_its_inputs():
    g, _ = _pair()
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    g.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 64, 64)
    with pytest.raises(svdag_py.SvdagError):
        g.trace_lighting(False, False)  # no colour pass yet
    g.trace_colors(np.ones((16, 3), np.float32))
    with pytest.raises(svdag_py.SvdagError):
        g.trace_lighting(True, False)  # no occlusion pass on this trace
    with pytest.raises(svdag_py.SvdagError):
        g.blur_ambient_occlusion()
    lit, _ = g.trace_lighting(False, False)
    assert lit.shape == (64 * 64,)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("integrate", [False, True])
def test_path_tracing_matches_oracle(integrate):
    # Paths follow the same PCG32 samples and exact voxel geometry on both
    # sides, so they take the same bounces; only powf's last bit in the
    # sRGB conversions may differ, which moves a channel by at most one count.
    g, o = _pair()
    for kind, mat, c, r in workloads.config1_edits():
        g.edit(kind, mat, c, r)
        o.edit(kind, mat, c, r)
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    W, H = 96, 64
    args = (cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], W, H)
    hits, _ = g.trace(*args)
    pal = np.random.default_rng(6).uniform(0.1, 0.9, (4, 3)).astype(np.float32)  # materials >= 4: checker
    col, _ = g.trace_colors(pal)
    img, n, _ = g.trace_path_tracing(pal, env=1.5, samples=2, depth=3, integrate=integrate, seed=7)
    oimg, oacc = o.path_trace(*args, oracle_py.colors(hits, pal), pal, env=1.5, samples=2, depth=3,
                              integrate=integrate, seed=7)
    img2, n2, _ = g.trace_path_tracing(pal, env=1.5, samples=1, depth=3, integrate=integrate, seed=8)
    oimg2, _ = o.path_trace(*args, oracle_py.colors(hits, pal), pal, env=1.5, samples=1, depth=3,
                            integrate=integrate, seed=8, accum=oacc, accumulated=2)
    assert (n, n2) == (2, 3)
    for a, b in ((img, oimg), (img2, oimg2)):
        d = np.abs(a.view(np.uint8).astype(np.int16) - b.view(np.uint8).astype(np.int16))
        assert d.max() <= 1 and (d.reshape(-1, 4).max(axis=1) > 0).mean() <= 0.01
    assert len(np.unique(img)) > 20  # lit, shadowed and sky pixels
    # A changed parameter restarts the accumulation, as the reference clears its buffer.
    _, n3, _ = g.trace_path_tracing(pal, env=1.0, samples=1, depth=3, integrate=integrate, seed=9)
    assert n3 == 1
    g.close()


@pytest.mark.gpu
def test_path_tracing_without_integration_needs_the_colour_pass():
    g, _ = _pair()
    cam = workloads.config1_camera(L10)
    rot = svdag_py.look_at(cam["position"], cam["target"])
    g.trace(cam["position"], rot, cam["fov"], cam["bmin"], cam["bmax"], 32, 32)
    with pytest.raises(svdag_py.SvdagError):
        g.trace_path_tracing(np.ones((2, 3), np.float32), integrate=False)
    img, n, _ = g.trace_path_tracing(np.ones((2, 3), np.float32), integrate=True)
    assert n == 1 and img.shape == (32 * 32,)
    g.close()
