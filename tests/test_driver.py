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


# ---- the driver against the FP64 oracle directly ---------------------------------------------------

TOTAL = np.array([2, 30, 26, 28])       # t, z, y, x of the whole volume
LOCAL = np.array([2, 24, 20, 22])       # this GPU's block
OFFSET = np.array([0, 3, 2, 4])         # its global offset


def _oracle_prob(ang, zs, views, mis=None):
    return oracle.Problem(ang[views], zs[views], D, PW, vol_dim=TOTAL[1:][::-1],
                          vol_start=OFFSET[1:][::-1], proj_dim=(NW, NH), mis=mis)


@pytest.mark.gpu
@pytest.mark.parametrize("mis", [None, [0.5, -0.7, 1.2, 0.03, -0.02, 0.04]])
def test_simple_project_matches_oracle(gpu, mis):
    # several chunks, two volumes (t), a sub-block of a larger volume, every second view
    # (projStride 2, projOffset 1): each selected view receives the sum over t of the oracle's
    # unit-step projection of the block, except in global column 0 (fastcone.project zeroes it in
    # every chunk window that starts there, so the caller's column 0 receives nothing); other
    # views are untouched
    ang, zs = _setup()
    rng = np.random.default_rng(2)
    vol = rng.standard_normal(LOCAL).astype(np.float32)
    base = rng.standard_normal((NV, NH, NW)).astype(np.float32)
    p = base.copy()
    mis32 = np.zeros(6, np.float32) if mis is None else np.asarray(mis, np.float32)
    driver.simpleProject(None, p, vol, ang, OFFSET, TOTAL, zs, D, chunk=(1, 10, 8, 11),
                         misAlignments=mis32, pixelWidth=PW, projStride=2, projOffset=1)
    sel = np.array([v for v in range(NV) if (v + 1) % 2 == 0])
    prob = _oracle_prob(ang.astype(np.float32).astype(np.float64),
                        zs.astype(np.float32).astype(np.float64), sel,
                        None if mis is None else mis32.astype(np.float64))
    ref = base.astype(np.float64).copy()
    add = sum(oracle.project(prob, vol[t].astype(np.float64), h=1.0) for t in range(LOCAL[0]))
    add[:, :, 0] = 0.0
    ref[sel] += add
    rest = np.setdiff1d(np.arange(NV), sel)
    assert np.array_equal(p[rest], base[rest])
    e = float(np.linalg.norm(p[sel] - ref[sel]) / np.linalg.norm(ref[sel]))
    assert e < 5e-5, e


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 1, 0])
def test_simple_backproject_matches_oracle(gpu, mode):
    # each t volume receives the oracle backprojection of the selected views with norm
    # 1 / (all views), the reference's backWeightings (mpi_fastcone.py:285)
    ang, zs = _setup()
    rng = np.random.default_rng(3)
    proj = rng.standard_normal((NV, NH, NW)).astype(np.float32)
    base = rng.standard_normal(LOCAL).astype(np.float32)
    v = base.copy()
    driver.simpleBackproject(None, proj, v, ang, OFFSET, TOTAL, zs, D, chunk=(1, 10, 8, 11),
                             kernelMode=mode, pixelWidth=PW, projStride=3, projOffset=0)
    sel = np.arange(0, NV, 3)
    prob = _oracle_prob(ang.astype(np.float32).astype(np.float64),
                        zs.astype(np.float32).astype(np.float64), sel)
    ref = oracle.backproject(prob, proj[sel].astype(np.float64), tuple(int(n) for n in LOCAL[1:]),
                             mode=mode,
                             norm=np.full(sel.size, 1.0 / NV))
    for t in range(LOCAL[0]):
        e = float(np.linalg.norm(v[t] - base[t] - ref) / np.linalg.norm(ref))
        assert e < 5e-5, (t, e)
