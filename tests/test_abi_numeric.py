"""The reference's C entry points called directly through ctypes with the reference's argument
layout (wrapper.c:112-249), against the FP64 oracle: GPU_CB_project and GPU_CB_backproject (the two
the Python surface does not reach), and blocks whose corners reach past the source radius, which
the reference projects and backprojects (its sweep start may be negative, coneBeamGPU.cu:1781).

Raw ABI semantics: projections / volume accumulate into the caller's arrays, norm scales the output
(FP) or the input (BP) per view, no column-0 zeroing (that is fastcone.project's), unit ray step.
"""
import ctypes

import numpy as np
import pytest

import oracle
from fastcone_amd import _native

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

_fp = ctypes.POINTER(ctypes.c_float)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _ip(a):
    return (ctypes.c_int * len(a))(*[int(x) for x in a])


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _args(c, vol, proj):
    """The 15 common arguments of the GPU_CB_* functions for case c."""
    nz, ny, nx = vol.shape
    vs = c["vs"]  # (z, y, x)
    return [ctypes.c_int(0), vol.ctypes.data_as(_fp), proj.ctypes.data_as(_fp),
            c["ang"].ctypes.data_as(_fp), c["z0"].ctypes.data_as(_fp), ctypes.c_float(c["D"]),
            ctypes.c_float(c["pw"]), _ip(c["dims"][::-1]), _ip(vs[::-1]),
            _ip((vs[2] + nx, vs[1] + ny, vs[0] + nz)),
            _ip((c["Nw"], c["Nh"], len(c["ang"]))), _ip((c["win"][1], c["win"][0])),
            c["woff"].ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            c["hoff"].ctypes.data_as(ctypes.POINTER(ctypes.c_int)), c["norm"].ctypes.data_as(_fp)]


def _case(rng, D=70.0, pw=0.8, nv=7, shape=(20, 22, 19), dims=(28, 30, 26), vs=(4, 2, 3),
          Nhw=(36, 40), win=(30, 33)):
    ang = _f(np.sort(rng.uniform(0, 2 * np.pi, nv)))
    z0 = _f(rng.uniform(dims[0] / 2 - 3, dims[0] / 2 + 3, nv))
    c = dict(ang=ang, z0=z0, D=float(D), pw=float(pw), dims=np.array(dims), vs=np.array(vs),
             Nh=Nhw[0], Nw=Nhw[1], win=win,
             woff=np.ascontiguousarray(rng.integers(-2, 6, nv), np.int32),
             hoff=np.ascontiguousarray(rng.integers(-2, 6, nv), np.int32),
             norm=_f(rng.uniform(0.5, 1.5, nv)), shape=shape)
    # the oracle in FP64 from the same FP32 inputs the C call receives
    c["prob"] = oracle.Problem(ang.astype(np.float64), z0.astype(np.float64),
                               float(np.float32(D)), float(np.float32(pw)),
                               vol_dim=dims[::-1], vol_start=vs[::-1], proj_dim=Nhw[::-1],
                               woff=c["woff"], hoff=c["hoff"])
    return c


def test_gpu_cb_project_accumulates_with_norm(gpu):
    L = _native.lib()
    rng = np.random.default_rng(90)
    c = _case(rng)
    vol = _f(rng.standard_normal(c["shape"]))
    base = _f(rng.standard_normal((len(c["ang"]),) + c["win"]))
    proj = base.copy()
    L.fc_clear_error()
    L.GPU_CB_project(*_args(c, vol, proj))
    _native.check()
    ref = base + oracle.project(c["prob"], vol, window=c["win"][::-1], h=1.0,
                                norm=c["norm"].astype(np.float64))
    assert rel(proj, ref) < 5e-5


@pytest.mark.parametrize("entry", ["GPU_CB_backproject", "GPU_CB_K_backproject"])
def test_gpu_cb_backproject_fdk_accumulates_with_norm(gpu, entry):
    L = _native.lib()
    rng = np.random.default_rng(91)
    c = _case(rng)
    y = _f(rng.standard_normal((len(c["ang"]),) + c["win"]))
    base = _f(rng.standard_normal(c["shape"]))
    vol = base.copy()
    L.fc_clear_error()
    if entry == "GPU_CB_backproject":
        L.GPU_CB_backproject(*_args(c, vol, y))
    else:
        L.GPU_CB_K_backproject(*_args(c, vol, y), ctypes.c_int(7))  # any mode but 0, 1 is FDK
    _native.check()
    ref = oracle.backproject(c["prob"], y, c["shape"], mode=2, norm=c["norm"].astype(np.float64))
    assert rel(vol - base, ref) < 5e-5


def test_block_reaching_past_the_source_radius(gpu):
    # source distance 9.3 voxels, block corners at radius ~12: some rays start inside the block
    # and, for every view, some voxels lie behind the source plane
    L = _native.lib()
    rng = np.random.default_rng(92)
    c = _case(rng, D=9.3, pw=0.7, nv=6, shape=(8, 16, 17), dims=(12, 18, 18), vs=(2, 1, 0),
              Nhw=(40, 44), win=(38, 42))
    vol = _f(rng.standard_normal(c["shape"]))
    proj = np.zeros((len(c["ang"]),) + c["win"], np.float32)
    L.fc_clear_error()
    L.GPU_CB_project(*_args(c, vol, proj))
    _native.check()
    ref = oracle.project(c["prob"], vol, window=c["win"][::-1], h=1.0,
                         norm=c["norm"].astype(np.float64))
    e_fp = rel(proj, ref)
    # backprojection: unweighted general geometry everywhere; FDK where the weight U^2 stays
    # moderate (|U| <= 2: away from the source plane, where FP32 positions lose accuracy)
    y = _f(rng.standard_normal((len(c["ang"]),) + c["win"]))
    v0 = np.zeros(c["shape"], np.float32)
    L.fc_clear_error()
    L.GPU_CB_K_backproject(*_args(c, v0, y), ctypes.c_int(0))
    _native.check()
    r0 = oracle.backproject(c["prob"], y, c["shape"], mode=0, norm=c["norm"].astype(np.float64))
    v2 = np.zeros(c["shape"], np.float32)
    L.GPU_CB_K_backproject(*_args(c, v2, y), ctypes.c_int(2))
    _native.check()
    r2 = oracle.backproject(c["prob"], y, c["shape"], mode=2, norm=c["norm"].astype(np.float64))
    o = c["prob"].origin
    zz, yy, xx = np.meshgrid(*(np.arange(n) for n in c["shape"]), indexing="ij")
    x, yv = xx + o[0], yy + o[1]
    umax = np.zeros(c["shape"])
    for th in c["ang"].astype(np.float64):
        umax = np.maximum(umax, np.abs(c["D"] / (c["D"] + x * np.cos(th) + yv * np.sin(th))))
    ok = umax <= 2.0
    behind = [(c["D"] + x * np.cos(th) + yv * np.sin(th) < 0).any() for th in c["ang"]]
    e_bp0, e_bp2 = rel(v0, r0), rel(v2[ok], r2[ok])
    print(f"block past the source: FP {e_fp:.2e}, BP mode 0 {e_bp0:.2e}, FDK (|U| <= 2, "
          f"{ok.mean():.0%} of voxels) {e_bp2:.2e}")
    assert all(behind) and ok.mean() > 0.2
    assert e_fp < 5e-5 and e_bp0 < 5e-5 and e_bp2 < 5e-5
