"""Property tests (hypothesis) over random small geometries.

CPU: linearity and exact adjointness of the FP64 oracle's projector pair, and the single-GPU
driver's chunk / window planning (chunks tile the block; every projected voxel of a chunk lies
inside its views' windows). GPU: linearity of the HIP projector, its dot-product test against
the HIP adjoint, and invariance of the host-memory ABI under random tiling limits.
"""
import math

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import oracle

geometries = st.fixed_dictionaries({
    "nv": st.integers(1, 5),
    "D": st.floats(25.0, 200.0),
    "pw": st.floats(0.5, 1.6),
    "shape": st.tuples(st.integers(3, 9), st.integers(3, 9), st.integers(3, 9)),
    "det": st.tuples(st.integers(4, 14), st.integers(4, 14)),
    "h": st.sampled_from([0.37, 0.5, 1.0]),
    "seed": st.integers(0, 2 ** 31 - 1),
})


def _problem(g):
    rng = np.random.default_rng(g["seed"])
    nz, ny, nx = g["shape"]
    ang = rng.uniform(0, 2 * math.pi, g["nv"])
    hmid = rng.uniform(nz / 2 - 1, nz / 2 + 1, g["nv"])
    nh, nw = g["det"]
    prob = oracle.Problem(ang, hmid, g["D"], g["pw"], vol_dim=(nx, ny, nz), proj_dim=(nw, nh))
    return prob, rng


@settings(max_examples=25, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(geometries)
def test_oracle_projector_is_linear_and_matched(g):
    prob, rng = _problem(g)
    x1, x2 = rng.standard_normal((2,) + g["shape"])
    a, b = rng.uniform(-2, 2, 2)
    p1 = oracle.project(prob, x1, h=g["h"])
    p2 = oracle.project(prob, x2, h=g["h"])
    p12 = oracle.project(prob, a * x1 + b * x2, h=g["h"])
    scale = max(np.abs(p1).max(), np.abs(p2).max(), 1e-30)
    assert np.abs(p12 - (a * p1 + b * p2)).max() <= 1e-12 * scale * (abs(a) + abs(b) + 1)
    y = rng.standard_normal(p1.shape)
    aty = oracle.adjoint(prob, y, g["shape"], h=g["h"])
    lhs, rhs = float(np.vdot(p1, y)), float(np.vdot(x1, aty))
    assert abs(lhs - rhs) <= 1e-11 * (np.abs(p1).sum() * np.abs(y).max() + 1e-30)


@settings(max_examples=40, deadline=None)
@given(st.tuples(st.integers(1, 3), st.integers(1, 30), st.integers(1, 30), st.integers(1, 30)),
       st.tuples(st.integers(1, 3), st.integers(1, 17), st.integers(1, 17), st.integers(1, 17)))
def test_driver_chunks_tile_the_block(dims, chunk):
    from fastcone_amd import driver
    plan = driver.ChunkPlan(dims, (0, 0, 0, 0), dims, chunk=chunk)
    cover = np.zeros(dims, int)
    for c in range(len(plan)):
        o, e = plan.local_offset[c], plan.extent[c]
        assert np.all(e >= 1)
        cover[o[0]:o[0] + e[0], o[1]:o[1] + e[1], o[2]:o[2] + e[2], o[3]:o[3] + e[3]] += 1
    assert np.all(cover == 1)


@settings(max_examples=20, deadline=None)
@given(st.integers(0, 2 ** 31 - 1), st.floats(30.0, 120.0), st.floats(0.6, 1.4))
def test_driver_windows_contain_projected_voxels(seed, D, pw):
    from fastcone_amd import driver
    rng = np.random.default_rng(seed)
    dims = np.array([1, 12, 14, 16])
    nv, nh, nw = 6, 24, 28
    ang = rng.uniform(0, 2 * math.pi, nv)
    zs = np.full(nv, dims[1] / 2.0) + rng.uniform(-2, 2, nv)
    plan = driver.ChunkPlan(dims, (0, 0, 0, 0), dims, chunk=(1, 7, 8, 9))
    wp = driver.WindowPlan(plan, ang, zs, D, (nv, nh, nw), pw)
    for c in range(len(plan)):
        o, e = plan.offset[c], plan.extent[c]
        pts = o[1:] + rng.uniform(0, 1, (64, 3)) * (e[1:] - 1) - dims[1:] // 2   # z, y, x
        for q, v in enumerate(wp.views[c]):
            ca, sa = math.cos(ang[v]), math.sin(ang[v])
            U = D / (D + pts[:, 2] * ca + pts[:, 1] * sa)
            w = (pts[:, 1] * ca - pts[:, 2] * sa) * U / pw + nw // 2
            hh = (pts[:, 0] - (zs[v] - dims[1] // 2)) * U / pw + nh // 2
            inside = (w >= 0) & (w <= nw - 1) & (hh >= 0) & (hh <= nh - 1)
            assert np.all(w[inside] >= wp.woff[c][q])
            assert np.all(w[inside] <= wp.woff[c][q] + wp.wsize[c][q] - 1)
            assert np.all(hh[inside] >= wp.hoff[c][q])
            assert np.all(hh[inside] <= wp.hoff[c][q] + wp.hsize[c][q] - 1)


@pytest.mark.gpu
@settings(max_examples=6, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(geometries)
def test_gpu_projector_linear_and_matched(gpu, g):
    import torch
    from fastcone_amd import ops
    from fastcone_amd.geometry import Geometry
    rng = np.random.default_rng(g["seed"])
    nz, ny, nx = g["shape"]
    ang = rng.uniform(0, 2 * math.pi, g["nv"])
    geom = Geometry.circular(ang, g["D"], g["pw"], g["shape"], g["det"])
    x1, x2 = (torch.from_numpy(rng.standard_normal(g["shape"]).astype(np.float32)).to(gpu)
              for _ in range(2))
    p1, p2 = ops.project(x1, geom, g["h"]), ops.project(x2, geom, g["h"])
    p12 = ops.project(2.0 * x1 - 0.5 * x2, geom, g["h"])
    err = float((p12 - (2.0 * p1 - 0.5 * p2)).abs().max())
    assert err <= 1e-5 * (float(p1.abs().max()) + float(p2.abs().max()) + 1e-30)
    y = torch.from_numpy(rng.standard_normal(geom.proj_shape).astype(np.float32)).to(gpu)
    aty = ops.adjoint(y, geom, g["h"])
    lhs = float((p1.double() * y.double()).sum())
    rhs = float((x1.double() * aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * float((p1.double().abs() * y.double().abs()).sum() + 1e-30)


@pytest.mark.gpu
@settings(max_examples=5, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(st.integers(0, 2 ** 31 - 1), st.integers(1, 6), st.integers(60, 2000))
def test_gpu_abi_tiling_invariance(gpu, seed, max_views, max_voxels):
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(seed)
    vol = rng.standard_normal((9, 10, 11)).astype(np.float32)
    nv = 7
    ang = np.sort(rng.uniform(0, 2 * math.pi, nv)).astype(np.float32)
    z0 = rng.uniform(3, 6, nv).astype(np.float32)
    p_whole = np.zeros((nv, 14, 16), np.float32)
    cone.project(vol, p_whole, ang, 60.0, z0, pixelWidth=0.9)
    v_whole = np.zeros_like(vol)
    cone.backproject(v_whole, p_whole, ang, 60.0, z0, pixelWidth=0.9)
    try:
        cone.set_tile_limits(max_voxels=max_voxels, max_views=max_views)
        p_t = np.zeros_like(p_whole)
        cone.project(vol, p_t, ang, 60.0, z0, pixelWidth=0.9)
        v_t = np.zeros_like(vol)
        cone.backproject(v_t, p_whole, ang, 60.0, z0, pixelWidth=0.9)
    finally:
        cone.set_tile_limits(0, 0)
    assert np.abs(p_t - p_whole).max() <= 1e-5 * (np.abs(p_whole).max() + 1e-30)
    assert np.abs(v_t - v_whole).max() <= 1e-5 * (np.abs(v_whole).max() + 1e-30)
