"""GPU parity of the FDK pre-processing kernel (fc_fdk_filter_dev, ops.fdk_filter) against the FP64
oracle: cosine weight from the oracle's own geometry vectors, then the reference's ramp
(fastcone.py:272-292) over a 2 Nw zero-padded row, divided by the pitch.

Tolerance: relative L2 1e-5 for FP32 output (FP32 FFT arithmetic, FP32 rsqrt for the weight);
16-bit outputs are held to their rounding (2^-8 / 2^-11 relative per element).
"""
import numpy as np
import pytest

import oracle
from fastcone_amd import _native

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(rng, nv=7, det=(26, 30), det_dim=(30, 34), mis=None, D=60.0, pw=0.9):
    from fastcone_amd.geometry import Geometry
    ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
    hmid = rng.uniform(9, 15, nv)
    woff = rng.integers(-3, 5, nv)
    hoff = rng.integers(-3, 5, nv)
    vol, vol_dim = (18, 20, 22), (24, 24, 26)
    g = Geometry.circular(ang, D, pw, vol, det, hmid=hmid, vol_dim=vol_dim, woff=woff, hoff=hoff,
                          misalign=mis, det_dim=det_dim)
    prob = oracle.Problem(ang, hmid, D, pw, vol_dim=vol_dim[::-1], proj_dim=det_dim[::-1],
                          woff=woff, hoff=hoff, mis=mis)
    return g, prob


@pytest.mark.parametrize("mis", [None, [0.6, -0.8, 1.5, 0.2, -0.25, 0.05]])
def test_fdk_filter_matches_oracle(gpu, mis):
    from fastcone_amd import ops
    rng = np.random.default_rng(70)
    g, prob = _case(rng, mis=mis)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    q = ops.fdk_filter(torch.from_numpy(y).to(gpu), g, 60.0, 0.9).cpu().numpy()
    ref = oracle.fdk_filter(prob, y, window=g.det_shape[::-1])
    e = rel(q, ref)
    print(f"fdk_filter vs oracle (mis={mis is not None}): rel L2 {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("din,dout", [("float32", "bfloat16"), ("float32", "float16"),
                                      ("bfloat16", "float32"), ("float16", "float32"),
                                      ("bfloat16", "bfloat16")])
def test_fdk_filter_storage_dtypes(gpu, din, dout):
    from fastcone_amd import ops
    rng = np.random.default_rng(71)
    g, prob = _case(rng)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    ti = torch.from_numpy(y).to(gpu).to(getattr(torch, din))
    q = ops.fdk_filter(ti, g, 60.0, 0.9, out_dtype=getattr(torch, dout))
    assert q.dtype == getattr(torch, dout)
    # the oracle sees the same rounded input
    ref = oracle.fdk_filter(prob, ti.float().cpu().numpy(), window=g.det_shape[::-1])
    e = rel(q.float().cpu().numpy(), ref)
    bound = {"float32": 1e-5, "float16": 1e-3, "bfloat16": 6e-3}[dout]
    print(f"fdk_filter {din} -> {dout}: rel L2 {e:.2e} (bound {bound})")
    assert e < bound


def test_fdk_filter_in_place_and_odd_row_count(gpu):
    # 3 views x 9 rows = 27 rows: the last row has no partner in its complex transform
    from fastcone_amd import ops
    rng = np.random.default_rng(72)
    g, prob = _case(rng, nv=3, det=(9, 21), det_dim=(12, 24))
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    t = torch.from_numpy(y).to(gpu)
    out = ops.fdk_filter(t, g, 60.0, 0.9, out=t)
    assert out.data_ptr() == t.data_ptr()
    ref = oracle.fdk_filter(prob, y, window=g.det_shape[::-1])
    assert rel(t.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("nw", [1, 2, 31, 32, 33, 64, 65, 127, 200, 513, 1024, 1025, 2049, 4096])
def test_ramp_every_fft_length(gpu, nw):
    # no cosine weight: the pure ramp (divided by pw) of the reference, FFT lengths 64 .. 8192
    from fastcone_amd import ops
    from fastcone_amd.geometry import Geometry
    rng = np.random.default_rng(73 + nw)
    g = Geometry.circular(np.array([0.1, 1.3, 2.0]), 80.0, 1.0, (4, 4, 4), (1, nw))
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    q = ops.fdk_filter(torch.from_numpy(y).to(gpu), g, 80.0, 0.8, cosine=False).cpu().numpy()
    ref = oracle.rampfilter(y) / 0.8
    e = rel(q, ref)
    assert e < 1e-5, (nw, e)


def test_fdk_filter_rejects_bad_arguments(gpu):
    from fastcone_amd import ops
    from fastcone_amd.geometry import Geometry
    g = Geometry.circular(np.array([0.1]), 80.0, 1.0, (4, 4, 4), (2, _native.FC_FILTER_MAX_COLS + 1))
    with pytest.raises(_native.FastconeError) as e:
        ops.fdk_filter(torch.zeros(g.proj_shape, device=gpu), g, 80.0, 1.0)
    assert e.value.code == _native.FC_ERR_ARGUMENT
    g0 = Geometry.circular(np.zeros(0), 80.0, 1.0, (4, 4, 4), (3, 5))
    q = ops.fdk_filter(torch.zeros(g0.proj_shape, device=gpu), g0, 80.0, 1.0)
    assert tuple(q.shape) == (0, 3, 5)
