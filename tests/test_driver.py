"""Single-GPU driver (the mpi_fastcone replacement): chunk and window planning on the CPU,
chunked projection / backprojection equal to the unchunked calls on the GPU."""
import numpy as np
import pytest

import oracle
from fastcone_amd import driver

DIMS = np.array([2, 24, 20, 22])       # t, z, y, x
NV, NH, NW = 9, 30, 34
D, PW = 60.0, 0.9


def _setup():
    ang = np.linspace(0, 2 * np.pi, NV, endpoint=False) + 0.2
    zs = np.full(NV, DIMS[1] / 2.0) + np.linspace(-2, 2, NV)
    return ang, zs


def test_chunk_plan_covers_the_block():
    plan = driver.ChunkPlan(DIMS, (0, 0, 0, 0), DIMS, chunk=(1, 10, 8, 22))
    cover = np.zeros(DIMS, int)
    for c in range(len(plan)):
        o, e = plan.local_offset[c], plan.extent[c]
        cover[o[0]:o[0] + e[0], o[1]:o[1] + e[1], o[2]:o[2] + e[2], o[3]:o[3] + e[3]] += 1
    assert np.all(cover == 1)


def test_windows_contain_every_projected_voxel_and_respect_stride():
    ang, zs = _setup()
    plan = driver.ChunkPlan(DIMS, (0, 0, 0, 0), DIMS, chunk=(1, 10, 8, 22))
    wp = driver.WindowPlan(plan, ang, zs, D, (NV, NH, NW), PW, proj_stride=2, proj_offset=1)
    prob = oracle.Problem(ang, zs, D, PW, vol_dim=DIMS[1:][::-1], proj_dim=(NW, NH))
    rng = np.random.default_rng(0)
    for c in range(len(plan)):
        assert all((v + 1) % 2 == 0 for v in wp.views[c])
        o, e = plan.offset[c], plan.extent[c]
        pts = o[1:] + rng.uniform(-1, 1, (200, 3)) * 0.999 + rng.uniform(0, 1, (200, 3)) * (e[1:] - 1)
        pts = pts - DIMS[1:] // 2  # z, y, x coordinates
        for q, v in enumerate(wp.views[c]):
            ca, sa = np.cos(ang[v]), np.sin(ang[v])
            U = D / (D + pts[:, 2] * ca + pts[:, 1] * sa)
            w = (pts[:, 1] * ca - pts[:, 2] * sa) * U / PW + NW // 2
            h = (pts[:, 0] - (zs[v] - DIMS[1] // 2)) * U / PW + NH // 2
            inside = (w >= 0) & (w <= NW - 1) & (h >= 0) & (h <= NH - 1)
            w, h = w[inside], h[inside]
            assert np.all(w >= wp.woff[c][q]) and np.all(w <= wp.woff[c][q] + wp.wsize[c][q] - 1)
            assert np.all(h >= wp.hoff[c][q]) and np.all(h <= wp.hoff[c][q] + wp.hsize[c][q] - 1)


def test_comm_must_be_single_rank():
    class Comm:
        size = 4
    with pytest.raises(ValueError):
        driver.simpleProject(Comm(), np.zeros((NV, NH, NW), np.float32),
                             np.zeros(DIMS, np.float32), *(_setup()[:1]), (0, 0, 0, 0), DIMS,
                             _setup()[1], D)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 0])
def test_chunked_equals_whole(gpu, mode):
    from fastcone_amd import fastcone as cone
    ang, zs = _setup()
    rng = np.random.default_rng(1)
    vol = rng.standard_normal(DIMS).astype(np.float32)
    chunk = (1, 10, 8, 22)
    p_drv = np.zeros((NV, NH, NW), np.float32)
    driver.simpleProject(None, p_drv, vol, ang, (0, 0, 0, 0), DIMS, zs, D, chunk=chunk,
                         pixelWidth=PW)
    p_ref = np.zeros_like(p_drv)
    for t in range(DIMS[0]):
        cone.project(np.ascontiguousarray(vol[t]), p_ref, ang, D, zs, pixelWidth=PW)
    # the direct call zeroes global column 0 (reference quirk); the driver's chunks do too
    assert np.linalg.norm(p_drv - p_ref) / np.linalg.norm(p_ref) < 1e-5
    v_drv = np.zeros(DIMS, np.float32)
    driver.simpleBackproject(None, p_ref, v_drv, ang, (0, 0, 0, 0), DIMS, zs, D, chunk=chunk,
                             kernelMode=mode, pixelWidth=PW)
    v_ref = np.zeros(DIMS[1:], np.float32)
    cone.backproject(v_ref, p_ref, ang, D, zs, kernelMode=mode, pixelWidth=PW)
    for t in range(DIMS[0]):
        assert np.linalg.norm(v_drv[t] - v_ref) / np.linalg.norm(v_ref) < 1e-5
