"""Shadow and ambient-occlusion passes on the GPU against the oracle
(Tracer::trace_shadows / trace_ambient_occlusion, tracer.cu:613-691)."""
import numpy as np
import pytest

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
