"""fc_fdk_filter_noalloc_dev (the FDK filter with its tables in a static device array instead of a
library workspace) against fc_fdk_filter_dev, through the C-ABI: two calls in a row, the second
with a smaller input (narrower rows, so a shorter FFT and different tables, and fewer views); the
results must be identical (same kernels, same FP64-built tables). Then the call inside a captured
HIP graph (where a device allocation would fail), replayed with eager work in between.
"""
import ctypes

import numpy as np
import pytest

from fastcone_amd import _native

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _call(name, x, out, geom, scale, stream=None):
    L = _native.lib()
    nh, nw = geom.det_shape
    shape = _native.int_array([nw, nh, geom.nviews])
    views = geom.views_on(x.device)
    s = torch.cuda.current_stream(x.device).cuda_stream if stream is None else stream
    rc = getattr(L, name)(x.device.index, x.data_ptr(), _native.FC_F32, out.data_ptr(),
                          _native.FC_F32, shape, views.data_ptr(), float(scale),
                          ctypes.c_void_p(int(s)))
    _native.check(rc)
    return views


def _geom(nv, nh, nw):
    from fastcone_amd.geometry import Geometry
    ang = np.linspace(0, 2 * np.pi, nv, endpoint=False) + 0.2
    return Geometry.circular(ang, 300.0, 0.8, (16, 16, 16), (nh, nw))


def test_noalloc_filter_equals_workspace_filter_on_two_calls(gpu):
    rng = np.random.default_rng(100)
    for nv, nh, nw in ((9, 40, 700), (4, 17, 130)):   # second call: smaller
        g = _geom(nv, nh, nw)
        x = torch.from_numpy(rng.standard_normal(g.proj_shape).astype(np.float32)).to(gpu)
        a = torch.empty_like(x)
        b = torch.full_like(x, float("nan"))
        _call("fc_fdk_filter_dev", x, a, g, 1.25)
        _call("fc_fdk_filter_noalloc_dev", x, b, g, 1.25)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (nv, nh, nw)
        assert float(a.abs().max()) > 0


def test_noalloc_filter_rejects_rows_beyond_the_tables(gpu):
    g = _geom(1, 2, _native.FC_FILTER_MAX_COLS + 1)
    x = torch.zeros(g.proj_shape, device=gpu)
    with pytest.raises(_native.FastconeError) as e:
        _call("fc_fdk_filter_noalloc_dev", x, x, g, 1.0)
    assert e.value.code == _native.FC_ERR_ARGUMENT


def test_noalloc_filter_under_graph_capture(gpu):
    rng = np.random.default_rng(101)
    g = _geom(6, 24, 300)
    x = torch.from_numpy(rng.standard_normal(g.proj_shape).astype(np.float32)).to(gpu)
    ref = torch.empty_like(x)
    _call("fc_fdk_filter_dev", x, ref, g, 1.0)
    out = torch.zeros_like(x)
    _call("fc_fdk_filter_noalloc_dev", x, out, g, 1.0)   # first call on the device: eager
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        views = g.views_on(gpu)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=s):
            _call("fc_fdk_filter_noalloc_dev", x, out, g, 1.0, stream=s.cuda_stream)
    for rep in range(3):
        out.fill_(float("nan"))
        # eager work on the same device between replays (the workspace filter on other data)
        y = torch.from_numpy(rng.standard_normal(g.proj_shape).astype(np.float32)).to(gpu)
        _call("fc_fdk_filter_dev", y, torch.empty_like(y), g, 2.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    del views
