"""CPU tests of the product's host logic: the C-ABI library loads and exports every declared
symbol, the FP64 geometry records agree with the oracle's restatement of coneBeamGPU.cu:1672-1734,
the projective matrices reproduce the reference's detector coordinates, and bad arguments are
reported through fc_last_error / Python exceptions (no GPU needed: validation precedes any device
call)."""
import ctypes
import subprocess

import numpy as np
import pytest

import oracle
from conftest import gpu_available
from fastcone_amd import _native, fastcone
from fastcone_amd.geometry import Geometry


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    names = _native.header_functions()
    assert "GPU_CB_K_backproject" in names and "fc_project_dev" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported
    # no CUDA-named shims are exported
    assert not [s for s in exported if s.startswith("cuda")]


def test_reference_entry_points_are_the_five_wrapper_symbols():
    names = [n for n in _native.header_functions() if n.startswith("GPU_CB_")]
    assert sorted(names) == sorted(["GPU_CB_K_backproject", "GPU_CB_misalign_backproject",
                                    "GPU_CB_backproject", "GPU_CB_misalign_project",
                                    "GPU_CB_project"])


@pytest.mark.parametrize("mis", [None, [0.7, -1.2, 2.0, 0.05, -0.08, 0.11]])
def test_view_records_match_oracle_geometry(mis):
    rng = np.random.default_rng(0)
    n = 9
    ang = rng.uniform(0, 2 * np.pi, n)
    hmid = rng.uniform(10, 20, n)
    woff = rng.integers(-5, 6, n)
    hoff = rng.integers(-5, 6, n)
    g = Geometry.circular(ang, 123.0, 0.75, (30, 32, 34), (40, 44), hmid=hmid,
                          vol_dim=(31, 32, 34), woff=woff, hoff=hoff, misalign=mis,
                          det_dim=(48, 50))
    src, det, ws, hs = oracle.geometry(ang, hmid, 123.0, 0.75, 31, mis)
    v = g.views
    assert np.allclose(v[:, 0:3], src, atol=1e-12)
    assert np.allclose(v[:, 6:9], ws, atol=1e-13)
    assert np.allclose(v[:, 9:12], hs, atol=1e-13)
    # window pixel (0,0) = det + (woff - Nw/2) ws + (hoff - Nh/2) hs
    d0 = det + (woff - 50 // 2)[:, None] * ws + (hoff - 48 // 2)[:, None] * hs
    assert np.allclose(v[:, 3:6], d0, atol=1e-11)


def test_projective_matrix_reproduces_reference_coordinates():
    # (a/c, b/c) must be the window coordinates of coneBeamGPU.cu:398-399 and 1/c the FDK U.
    rng = np.random.default_rng(1)
    n = 6
    ang = rng.uniform(0, 2 * np.pi, n)
    hmid = rng.uniform(10, 20, n)
    Nw, Nh, Nz = 50, 48, 31
    woff = rng.integers(-5, 6, n)
    hoff = rng.integers(-5, 6, n)
    D, pw = 123.0, 0.75
    g = Geometry.circular(ang, D, pw, (Nz, 32, 34), (40, 44), hmid=hmid, woff=woff, hoff=hoff,
                          det_dim=(Nh, Nw))
    X = rng.uniform(-15, 15, (100, 3))
    for t in range(n):
        M = g.views[t, 12:24].reshape(3, 4)
        abc = X @ M[:, :3].T + M[:, 3]
        u, v, c = abc[:, 0] / abc[:, 2], abc[:, 1] / abc[:, 2], abc[:, 2]
        ca, sa = np.cos(ang[t]), np.sin(ang[t])
        U = D / (D + X[:, 0] * ca + X[:, 1] * sa)
        w_ref = (X[:, 1] * ca - X[:, 0] * sa) * U / pw + Nw // 2 - woff[t]
        h_ref = (X[:, 2] - (hmid[t] - Nz // 2)) * U / pw + Nh // 2 - hoff[t]
        assert np.allclose(u, w_ref, atol=1e-9)
        assert np.allclose(v, h_ref, atol=1e-9)
        assert np.allclose(1.0 / c, U, atol=1e-12)


def test_projective_matrix_misaligned_matches_ray_plane_intersection():
    rng = np.random.default_rng(2)
    n = 4
    mis = [0.7, -1.2, 2.0, 0.05, -0.08, 0.11]
    ang = rng.uniform(0, 2 * np.pi, n)
    g = Geometry.circular(ang, 80.0, 0.9, (20, 20, 20), (30, 30), misalign=mis)
    src, det, ws, hs = oracle.geometry(ang, np.full(n, 10.0), 80.0, 0.9, 20, mis)
    X = rng.uniform(-8, 8, (50, 3))
    for t in range(n):
        M = g.views[t, 12:24].reshape(3, 4)
        abc = X @ M[:, :3].T + M[:, 3]
        u, v = abc[:, 0] / abc[:, 2], abc[:, 1] / abc[:, 2]
        # coneBeamGPU.cu:235-243 with window pixel (0,0) at global (0,0)
        npl = np.cross(ws[t], hs[t])
        dist = ((det[t] - src[t]) @ npl) / ((X - src[t]) @ npl)
        I = dist[:, None] * (X - src[t]) + src[t] - det[t]
        wdw, wdh, hdh = ws[t] @ ws[t], ws[t] @ hs[t], hs[t] @ hs[t]
        idw, idh = I @ ws[t], I @ hs[t]
        dd = wdw * hdh - wdh * wdh
        assert np.allclose(u, (idw * hdh - idh * wdh) / dd + 15, atol=1e-9)
        assert np.allclose(v, (idh * wdw - idw * wdh) / dd + 15, atol=1e-9)


def _abi_args(nv=3, vol=(8, 8, 8), win=(10, 10)):
    v = np.zeros(vol[::-1], np.float32)
    p = np.zeros((nv, win[1], win[0]), np.float32)
    ang = np.linspace(0, 1, nv).astype(np.float32)
    hm = np.full(nv, 4.0, np.float32)
    return v, p, ang, hm


def test_abi_reports_bad_arguments_without_touching_memory():
    L = _native.lib()
    v, p, ang, hm = _abi_args()
    ip = lambda a: (ctypes.c_int * len(a))(*a)
    woff = ip([0, 0, 0])
    hoff = ip([0, 0, 0])
    norm = (ctypes.c_float * 3)(1, 1, 1)
    fp = ctypes.POINTER(ctypes.c_float)
    base = [ctypes.c_int(0), v.ctypes.data_as(fp), p.ctypes.data_as(fp), ang.ctypes.data_as(fp),
            hm.ctypes.data_as(fp), ctypes.c_float(50.0), ctypes.c_float(1.0)]
    # empty volume block (volEnd == volStart)
    L.fc_clear_error()
    L.GPU_CB_project(*base, ip([8, 8, 8]), ip([0, 0, 0]), ip([0, 8, 8]), ip([10, 10, 3]),
                     ip([10, 10]), woff, hoff, norm)
    assert L.fc_last_error() == _native.FC_ERR_ARGUMENT
    assert b"empty volume block" in L.fc_last_error_message()
    # non-positive source distance
    base[5] = ctypes.c_float(0.0)
    L.fc_clear_error()
    L.GPU_CB_K_backproject(*base, ip([8, 8, 8]), ip([0, 0, 0]), ip([8, 8, 8]), ip([10, 10, 3]),
                           ip([10, 10]), woff, hoff, norm, ctypes.c_int(2))
    assert L.fc_last_error() == _native.FC_ERR_ARGUMENT
    assert b"sourceSampleDistance" in L.fc_last_error_message()
    assert np.all(v == 0) and np.all(p == 0)


@pytest.mark.skipif(gpu_available(), reason="checks the validation order on a host without GPU")
def test_abi_accepts_blocks_that_reach_the_source():
    # The reference projects and backprojects blocks whose corners lie beyond the source radius
    # (its sweep start may go negative, coneBeamGPU.cu:1781): validation must pass them, so on a
    # host without a GPU the only error left is the missing device.
    L = _native.lib()
    v, p, ang, hm = _abi_args()
    ip = lambda a: (ctypes.c_int * len(a))(*a)
    woff, hoff = ip([0, 0, 0]), ip([0, 0, 0])
    norm = (ctypes.c_float * 3)(1, 1, 1)
    fp = ctypes.POINTER(ctypes.c_float)
    base = [ctypes.c_int(0), v.ctypes.data_as(fp), p.ctypes.data_as(fp), ang.ctypes.data_as(fp),
            hm.ctypes.data_as(fp), ctypes.c_float(3.0), ctypes.c_float(1.0)]
    dims = [ip([8, 8, 8]), ip([0, 0, 0]), ip([8, 8, 8]), ip([10, 10, 3]), ip([10, 10]), woff,
            hoff, norm]
    for call in (lambda: L.GPU_CB_project(*base, *dims),
                 lambda: L.GPU_CB_K_backproject(*base, *dims, ctypes.c_int(2)),
                 lambda: L.GPU_CB_misalign_project(*base, *dims, (ctypes.c_float * 6)())):
        L.fc_clear_error()
        call()
        assert L.fc_last_error() == _native.FC_ERR_DEVICE, L.fc_last_error_message()
    assert np.all(v == 0) and np.all(p == 0)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_abi_reports_missing_device():
    L = _native.lib()
    v, p, ang, hm = _abi_args()
    with pytest.raises(_native.FastconeError) as e:
        fastcone.project(v, p, ang, 50.0, hm)
    assert e.value.code == _native.FC_ERR_DEVICE


def test_python_surface_validation():
    v, p, ang, hm = _abi_args()
    with pytest.raises(TypeError):
        fastcone.project(v.astype(np.float64), p, ang, 50.0, hm)
    with pytest.raises(TypeError):
        fastcone.backproject(v, p.astype(np.float64), ang, 50.0, hm)
    with pytest.raises(ValueError):
        fastcone.project(v, p, ang[:2], 50.0, hm)
    with pytest.raises(ValueError):
        fastcone.backproject(v, p, ang, 50.0, hm[:1])
    with pytest.raises(ValueError):
        fastcone.backproject(v, p, ang, 50.0, hm, kernelMode=2, misAlignments=np.zeros(6))
    with pytest.raises(ValueError):
        fastcone.project(v, p[:, :, ::2], ang, 50.0, hm)


def test_rampfilter_matches_oracle():
    rng = np.random.default_rng(3)
    p = rng.standard_normal((5, 7, 33)).astype(np.float32)
    expect = oracle.rampfilter(p)
    q = p.copy()
    fastcone.rampfilter(q)
    assert np.abs(q - expect).max() < 1e-6 * np.abs(expect).max()


def test_ramp_response_matches_oracle():
    for nw in (8, 33, 768):
        assert np.allclose(fastcone.ramp_response(nw), oracle.ramp_filter_response(nw),
                           rtol=1e-15, atol=0)


def test_gaussrampfilter_window_matches_scipy_gaussian():
    # the window is a Gaussian over frequency indices with standard deviation scale * Nw (the
    # reference builds it with scipy's gaussian window centred on the DC bin)
    from scipy.signal import windows
    nw, scale = 64, 0.125
    ratio = fastcone.ramp_response(nw, "gauss", scale) / fastcone.ramp_response(nw)
    g = windows.gaussian(nw + 1, scale * nw)  # odd length: index nw // 2 is the centre
    for k in range(nw // 2 + 1):
        assert abs(ratio[2 * k] - g[nw // 2 + k]) < 1e-12  # padded bin 2k is frequency k / Nw


def test_windowed_ramp_filters():
    rng = np.random.default_rng(4)
    p = rng.standard_normal((3, 4, 40)).astype(np.float32)
    wide, plain = p.copy(), p.copy()
    fastcone.gaussrampfilter(wide, scale=1e6)  # window -> 1
    fastcone.rampfilter(plain)
    assert np.allclose(wide, plain, rtol=1e-6, atol=1e-6)
    H = fastcone.ramp_response(40, "cos")
    assert abs(H[-1]) < 1e-15 and np.all(H <= fastcone.ramp_response(40) + 1e-15)
    c = p.copy()
    fastcone.cosrampfilter(c)
    # every view is filtered, and the cosine window removes energy near the Nyquist frequency
    hi = lambda a: np.abs(np.fft.rfft(a.astype(np.float64), axis=2)[..., -6:]).sum()
    assert np.all([np.abs(c[t] - p[t]).max() > 0 for t in range(3)])
    assert hi(c) < 0.5 * hi(plain)
    with pytest.raises(ValueError):
        fastcone.ramp_response(8, "hann")
