"""GPU parity: the HIP kernels (through the C-ABI) against the FP64 oracle on the same inputs.

Tolerances (relative L2 = |gpu - oracle| / |oracle|):
  * FP32 storage: 5e-5 on N(0,1) volumes (FP32 positions and interpolation weights on rough
    data), 1e-4 on the configuration-0 phantom (the workload target);
  * BF16 / FP16 storage: the oracle is fed the same rounded input, so the bound is again the FP32
    arithmetic one (5e-5); the error against the unrounded input is reported.
"""
import math
import os

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


def _setup(rng, n=12, vol=(18, 20, 22), vol_dim=(24, 24, 26), vol_start=(2, 3, 1),
           det=(26, 30), det_dim=(30, 34), D=60.0, pw=0.9, mis=None, ang=None):
    """C-order vol=(nz,ny,nx), vol_dim=(Nz,Ny,Nx), vol_start=(z,y,x), det=(nh,nw)."""
    from fastcone_amd.geometry import Geometry
    if ang is None:
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    n = len(ang)
    hmid = rng.uniform(vol_dim[0] / 2 - 3, vol_dim[0] / 2 + 3, n)
    woff = rng.integers(-3, 5, n)
    hoff = rng.integers(-3, 5, n)
    g = Geometry.circular(ang, D, pw, vol, det, hmid=hmid, vol_dim=vol_dim, vol_start=vol_start,
                          woff=woff, hoff=hoff, misalign=mis, det_dim=det_dim)
    prob = oracle.Problem(ang, hmid, D, pw, vol_dim=vol_dim[::-1], vol_start=vol_start[::-1],
                          proj_dim=det_dim[::-1], woff=woff, hoff=hoff, mis=mis)
    return g, prob


def _round(x, dtype):
    return torch.from_numpy(np.asarray(x, np.float32)).to(dtype).float().numpy().astype(np.float64)


@pytest.mark.parametrize("h", [1.0, 0.5])
@pytest.mark.parametrize("mis", [None, [0.6, -0.8, 1.5, 0.03, -0.04, 0.05]])
def test_project_f32_parity(gpu, h, mis):
    from fastcone_amd import ops
    rng = np.random.default_rng(11)
    g, prob = _setup(rng, mis=mis)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    norm = rng.uniform(0.5, 1.5, g.nviews).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu), g, step=h, norm=torch.from_numpy(norm)).cpu()
    ref = oracle.project(prob, x, window=g.det_shape[::-1], h=h, norm=norm)
    assert rel(p, ref) < 5e-5


@pytest.mark.parametrize("layout,cap,quad", [("octet", None, None), ("plain", None, None),
                                             ("lds", None, None), ("lds", None, "always"),
                                             ("lds", None, "never"), ("lds", "600", "always"),
                                             ("lds", "600", "never"), ("lds", "0", None)])
def test_project_layouts_agree(gpu, layout, cap, quad, monkeypatch):
    # cap lowers the LDS brick budget so that slabs hold plain values instead of (value,
    # difference) pairs and some (600) or all (0) slabs take the global fallback; quad forces
    # every quad of views into shared (one brick per 16x8 half tile for four views) or single
    # (one view's whole tile per block) mode
    from fastcone_amd import ops
    monkeypatch.setenv("FASTCONE_FP_LAYOUT", layout)
    if cap is not None:
        monkeypatch.setenv("FASTCONE_FP_LDS_CAP", cap)
    if quad is not None:
        monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    rng = np.random.default_rng(12)
    g, prob = _setup(rng)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu), g, step=0.5).cpu()
    ref = oracle.project(prob, x, window=g.det_shape[::-1], h=0.5)
    assert rel(p, ref) < 5e-5


@pytest.mark.parametrize("quad", [None, "always", "never"])
@pytest.mark.parametrize("seed", [14, 33])
def test_project_lds_steep_and_wide_cone(gpu, seed, quad, monkeypatch):
    # short source distance and a tilted detector: large cone angles, views whose dominant axis
    # is z for some tiles, and bricks that exceed the LDS budget; seed 33 has a tile whose corner
    # ray (missing the volume) points against the tile's dominant axis
    from fastcone_amd import ops
    if quad is not None:
        monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    rng = np.random.default_rng(seed)
    g, prob = _setup(rng, D=28.0, pw=1.6, mis=[0.0, 0.0, 0.0, 0.9, 0.3, 0.2])
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu), g, step=0.5).cpu()
    ref = oracle.project(prob, x, window=g.det_shape[::-1], h=0.5)
    assert rel(p, ref) < 5e-5


@pytest.mark.parametrize("quad", [None, "always", "never"])
@pytest.mark.parametrize("nviews", [13, 16])
def test_project_quads_dense_views(gpu, nviews, quad, monkeypatch):
    # consecutive views 0.5 deg apart share their bricks (shared mode) unless a mode is forced;
    # 13 views leave a quad with three padding views; the views straddle the 45 deg diagonal,
    # where a tile's dominant axis can change within a quad
    from fastcone_amd import ops
    if quad is not None:
        monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    rng = np.random.default_rng(40 + nviews)
    ang = np.deg2rad(41.0 + 0.5 * np.arange(nviews))
    g, prob = _setup(rng, ang=ang, vol=(20, 24, 22), vol_dim=(26, 28, 26), det=(40, 44),
                     det_dim=(44, 48))
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    norm = rng.uniform(0.5, 1.5, g.nviews).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu), g, step=0.5, norm=torch.from_numpy(norm)).cpu()
    ref = oracle.project(prob, x, window=g.det_shape[::-1], h=0.5, norm=norm)
    assert rel(p, ref) < 5e-5


@pytest.mark.parametrize("quad", [None, "always", "never"])
@pytest.mark.parametrize("nviews", [1, 2, 3, 5])
def test_project_small_view_counts(gpu, nviews, quad, monkeypatch):
    # stacks that leave one to three padding views in their last quad of four
    from fastcone_amd import ops
    if quad is not None:
        monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    rng = np.random.default_rng(60 + nviews)
    ang = np.deg2rad(100.0 + 0.5 * np.arange(nviews))
    g, prob = _setup(rng, ang=ang)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu), g, step=0.5).cpu()
    ref = oracle.project(prob, x, window=g.det_shape[::-1], h=0.5)
    assert rel(p, ref) < 5e-5


@pytest.mark.parametrize("quad", ["always", "never"])
@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
def test_project_16bit_storage_dense_views(gpu, dtype, quad, monkeypatch):
    # 16-bit bricks filled for four views at once (shared quads) and one view at a time
    from fastcone_amd import ops
    monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    dt = getattr(torch, dtype)
    rng = np.random.default_rng(50)
    ang = np.deg2rad(10.0 + 0.5 * np.arange(8))
    g, prob = _setup(rng, ang=ang, vol=(20, 24, 22), vol_dim=(26, 28, 26), det=(40, 44),
                     det_dim=(44, 48))
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu).to(dt), g, step=0.5).cpu().numpy()
    ref_rounded = oracle.project(prob, _round(x, dt), window=g.det_shape[::-1], h=0.5)
    assert rel(p, ref_rounded) < 5e-5


@pytest.mark.parametrize("dtype", ["float32", "bfloat16"])
@pytest.mark.parametrize("quad", ["always", "never"])
def test_project_pair_bricks_bit_identical(gpu, dtype, quad, monkeypatch):
    # A pair brick holds f(d) and f(d + 1) - f(d), the subtraction done once at the fill, and
    # interpolates with fmaf(a, f1 - f0, f0): the operations lerp(f0, f1, a) performs on plain
    # values, in the same order; the cell reuse only skips rereading the same four pairs. So the
    # projections equal those of plain-value bricks (FASTCONE_FP_PAIRS=0) bit for bit, for x- and
    # y-dominant views (angles near 10 and 100 degrees), shared quads and single views.
    from fastcone_amd import ops
    monkeypatch.setenv("FASTCONE_FP_LAYOUT", "lds")
    monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    dt = getattr(torch, dtype)
    rng = np.random.default_rng(51)
    ang = np.deg2rad(np.concatenate([10.0 + 0.5 * np.arange(8), 100.0 + 0.5 * np.arange(8)]))
    g, _ = _setup(rng, ang=ang, vol=(20, 24, 22), vol_dim=(26, 28, 26), det=(40, 44),
                  det_dim=(44, 48))
    x = torch.from_numpy(rng.standard_normal(g.vol_shape).astype(np.float32)).to(gpu).to(dt)
    p_pairs = ops.project(x, g, step=0.5)
    monkeypatch.setenv("FASTCONE_FP_PAIRS", "0")
    p_plain = ops.project(x, g, step=0.5)
    torch.cuda.synchronize()
    assert float(p_pairs.abs().max()) > 0.0
    assert torch.equal(p_pairs, p_plain)


@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
def test_project_16bit_storage(gpu, dtype):
    from fastcone_amd import ops
    dt = getattr(torch, dtype)
    rng = np.random.default_rng(13)
    g, prob = _setup(rng)
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    xs = torch.from_numpy(x).to(gpu).to(dt)
    p = ops.project(xs, g, step=0.5).cpu().numpy()
    ref_rounded = oracle.project(prob, _round(x, dt), window=g.det_shape[::-1], h=0.5)
    ref = oracle.project(prob, x, window=g.det_shape[::-1], h=0.5)
    assert rel(p, ref_rounded) < 5e-5
    print(f"{dtype} storage: rel L2 vs FP64 oracle of unrounded input = {rel(p, ref):.3e}")
    assert rel(p, ref) < (1e-2 if dtype == "bfloat16" else 2e-3)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_backproject_f32_parity(gpu, mode):
    from fastcone_amd import ops
    rng = np.random.default_rng(20 + mode)
    g, prob = _setup(rng)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    norm = rng.uniform(0.5, 1.5, g.nviews).astype(np.float32)
    v = ops.backproject(torch.from_numpy(y).to(gpu), g, mode=mode,
                        norm=torch.from_numpy(norm)).cpu()
    ref = oracle.backproject(prob, y, g.vol_shape, mode=mode, norm=norm)
    assert rel(v, ref) < 5e-5


def test_backproject_misaligned_mode0(gpu):
    from fastcone_amd import ops
    rng = np.random.default_rng(24)
    g, prob = _setup(rng, mis=[0.6, -0.8, 1.5, 0.03, -0.04, 0.05])
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    v = ops.backproject(torch.from_numpy(y).to(gpu), g, mode=0).cpu()
    ref = oracle.backproject(prob, y, g.vol_shape, mode=0, norm=np.ones(g.nviews))
    assert rel(v, ref) < 5e-5


@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
def test_backproject_16bit_storage(gpu, dtype):
    from fastcone_amd import ops
    dt = getattr(torch, dtype)
    rng = np.random.default_rng(25)
    g, prob = _setup(rng)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    v = ops.backproject(torch.from_numpy(y).to(gpu).to(dt), g, mode=2).cpu().numpy()
    ref_rounded = oracle.backproject(prob, _round(y, dt), g.vol_shape, mode=2,
                                     norm=np.ones(g.nviews))
    assert rel(v, ref_rounded) < 5e-5


def test_backproject_accumulates_into_out(gpu):
    from fastcone_amd import ops
    rng = np.random.default_rng(26)
    g, prob = _setup(rng)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    base = rng.standard_normal(g.vol_shape).astype(np.float32)
    out = torch.from_numpy(base.copy()).to(gpu)
    ops.backproject(torch.from_numpy(y).to(gpu), g, mode=2, out=out)
    ref = oracle.backproject(prob, y, g.vol_shape, mode=2, norm=np.ones(g.nviews))
    assert rel(out.cpu().numpy() - base, ref) < 5e-5


@pytest.mark.parametrize("nviews", [1, 2])
@pytest.mark.parametrize("dtype", ["float32", "bfloat16"])
def test_backproject_one_and_two_views(gpu, nviews, dtype):
    # the view loop's first and last views: the next view's window word and rows are read only
    # while a next view exists
    from fastcone_amd import ops
    dt = getattr(torch, dtype)
    rng = np.random.default_rng(27 + nviews)
    g, prob = _setup(rng, ang=np.sort(rng.uniform(0, 2 * np.pi, nviews)))
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    norm = rng.uniform(0.5, 1.5, nviews).astype(np.float32)
    v = ops.backproject(torch.from_numpy(y).to(gpu).to(dt), g, mode=2,
                        norm=torch.from_numpy(norm)).cpu().numpy()
    ref = oracle.backproject(prob, _round(y, dt), g.vol_shape, mode=2, norm=norm)
    assert float(np.abs(ref).max()) > 0.0
    assert rel(v, ref) < 5e-5


@pytest.mark.parametrize("layout", ["lds", "plain"])
def test_adjoint_paths_agree(gpu, layout, monkeypatch):
    # lds: voxel-owned bricks with LDS accumulation (default); plain: global-atomic scatter
    from fastcone_amd import ops
    monkeypatch.setenv("FASTCONE_FP_LAYOUT", layout)
    rng = np.random.default_rng(32)
    g, prob = _setup(rng)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    aty = ops.adjoint(torch.from_numpy(y).to(gpu), g, step=0.5).cpu().numpy()
    ref = oracle.adjoint(prob, y, g.vol_shape, h=0.5)
    assert rel(aty, ref) < 5e-5


def test_adjoint_steep_and_wide_cone(gpu):
    from fastcone_amd import ops
    rng = np.random.default_rng(33)
    g, prob = _setup(rng, D=28.0, pw=1.6, mis=[0.0, 0.0, 0.0, 0.9, 0.3, 0.2])
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    aty = ops.adjoint(torch.from_numpy(y).to(gpu), g, step=0.5).cpu().numpy()
    ref = oracle.adjoint(prob, y, g.vol_shape, h=0.5)
    assert rel(aty, ref) < 5e-5


# 0.3: runs of up to 6 same-cell samples merged without a count; 0.2: runs capped at 4
@pytest.mark.parametrize("h", [1.0, 0.5, 0.3, 0.2])
def test_adjoint_parity_and_dot_product(gpu, h):
    from fastcone_amd import ops
    rng = np.random.default_rng(30)
    g, prob = _setup(rng, mis=[0.2, 0.1, -0.4, 0.01, 0.02, -0.01])
    x = rng.standard_normal(g.vol_shape).astype(np.float32)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    yt = torch.from_numpy(y).to(gpu)
    aty = ops.adjoint(yt, g, step=h).cpu().numpy()
    ref = oracle.adjoint(prob, y, g.vol_shape, h=h)
    assert rel(aty, ref) < 5e-5
    ax = ops.project(torch.from_numpy(x).to(gpu), g, step=h).cpu().numpy()
    lhs = np.vdot(ax.astype(np.float64), y.astype(np.float64))
    rhs = np.vdot(x.astype(np.float64), aty.astype(np.float64))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_step_too_small_for_the_volume_is_rejected(gpu):
    # the sample loops count in FP32, exact below 2^24: a 16384 x 3700 slice (the largest
    # dimension allowed, diagonal 16.8k voxels) at step 1e-3 could need 1.68e7 samples per ray,
    # so the projector and the adjoint refuse it instead of rounding k
    from fastcone_amd import ops
    from fastcone_amd._native import FastconeError
    from fastcone_amd.geometry import Geometry
    g = Geometry.circular(np.array([0.3]), 60000.0, 1.0, (1, 3700, 16384), (8, 8))
    x = torch.zeros(g.vol_shape, device=gpu)
    with pytest.raises(FastconeError, match="2\\^24"):
        ops.project(x, g, step=1e-3)
    y = torch.zeros(g.proj_shape, device=gpu)
    with pytest.raises(FastconeError, match="2\\^24"):
        ops.adjoint(y, g, step=1e-3)
    # the same geometry at step 0.5 runs
    ops.project(x, g, step=0.5)
    torch.cuda.synchronize()


def test_sample_count_exact(gpu):
    from fastcone_amd import ops
    rng = np.random.default_rng(31)
    g, prob = _setup(rng)
    c = ops.sample_count(g, 0.5, gpu).cpu().numpy()
    ref = oracle.sample_count(prob, g.vol_shape, window=g.det_shape[::-1], h=0.5)
    assert np.array_equal(c, ref)


def test_convert_round_to_nearest_even(gpu):
    from fastcone_amd import ops
    x = torch.randn(100000, device=gpu) * 100
    for dt in (torch.bfloat16, torch.float16):
        assert torch.equal(ops.convert(x, dt), x.to(dt))
        assert torch.equal(ops.convert(x.to(dt), torch.float32), x.to(dt).float())


# ---- the reference ABI (host memory) -------------------------------------------------------------

def _abi_case(rng, nv=10):
    Nz, Ny, Nx = 30, 28, 26
    vs = np.array([4, 2, 3])          # (z, y, x)
    shape = (20, 22, 19)
    vol = rng.standard_normal(shape).astype(np.float32)
    Nh, Nw = 36, 40
    win = (30, 33)                    # (h, w)
    ang = np.sort(rng.uniform(0, 2 * np.pi, nv)).astype(np.float32)
    z0 = rng.uniform(12, 18, nv).astype(np.float32)
    woff = rng.integers(-2, 6, nv).astype(np.int32)
    hoff = rng.integers(-2, 6, nv).astype(np.int32)
    prob = oracle.Problem(ang.astype(np.float64), z0.astype(np.float64), 70.0, 0.8,
                          vol_dim=(Nx, Ny, Nz), vol_start=vs[::-1], proj_dim=(Nw, Nh),
                          woff=woff, hoff=hoff)
    return dict(vol=vol, vs=vs, dims=np.array([Nz, Ny, Nx]), win=win, ang=ang, z0=z0,
                woff=woff, hoff=hoff, pdims=np.array([nv, Nh, Nw]), prob=prob)


def test_abi_project_matches_oracle_with_reference_conventions(gpu):
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(40)
    c = _abi_case(rng)
    base = rng.standard_normal((len(c["ang"]),) + c["win"]).astype(np.float32)
    proj = base.copy()
    cone.project(c["vol"], proj, c["ang"], 70.0, c["z0"], volStart=c["vs"], volDims=c["dims"],
                 projWOffset=c["woff"], projHOffset=c["hoff"], projDims=c["pdims"],
                 pixelWidth=0.8)
    ref = oracle.project(c["prob"], c["vol"], window=c["win"][::-1], h=1.0) + base
    for t in range(len(c["ang"])):   # the reference zeroes up to global column 0
        if c["woff"][t] <= 0:
            ref[t, :, :1 + c["woff"][t]] = 0.0
    assert rel(proj, ref) < 5e-5


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_abi_backproject_matches_oracle(gpu, mode):
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(41 + mode)
    c = _abi_case(rng)
    y = rng.standard_normal((len(c["ang"]),) + c["win"]).astype(np.float32)
    base = rng.standard_normal(c["vol"].shape).astype(np.float32)
    vol = base.copy()
    cone.backproject(vol, y, c["ang"], 70.0, c["z0"], volStart=c["vs"], volDims=c["dims"],
                     projWOffset=c["woff"], projHOffset=c["hoff"], projDims=c["pdims"],
                     kernelMode=mode, pixelWidth=0.8)
    ref = oracle.backproject(c["prob"], y, vol.shape, mode=mode,
                             norm=np.full(len(c["ang"]), 1.0 / len(c["ang"])))
    assert rel(vol - base, ref) < 5e-5


def test_abi_misaligned_backproject(gpu):
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(45)
    c = _abi_case(rng)
    mis = np.array([0.6, -0.8, 1.5, 0.03, -0.04, 0.05], np.float32)
    c["prob"] = oracle.Problem(c["ang"].astype(np.float64), c["z0"].astype(np.float64), 70.0, 0.8,
                               vol_dim=c["dims"][::-1], vol_start=c["vs"][::-1],
                               proj_dim=c["pdims"][1:][::-1], woff=c["woff"], hoff=c["hoff"],
                               mis=mis.astype(np.float64))
    y = rng.standard_normal((len(c["ang"]),) + c["win"]).astype(np.float32)
    vol = np.zeros_like(c["vol"])
    cone.backproject(vol, y, c["ang"], 70.0, c["z0"], volStart=c["vs"], volDims=c["dims"],
                     projWOffset=c["woff"], projHOffset=c["hoff"], projDims=c["pdims"],
                     kernelMode=0, misAlignments=mis, pixelWidth=0.8)
    ref = oracle.backproject(c["prob"], y, vol.shape, mode=0,
                             norm=np.full(len(c["ang"]), 1.0 / len(c["ang"])))
    assert rel(vol, ref) < 5e-5


def test_abi_fbp_matches_oracle_pipeline(gpu):
    # fastcone.fbp = the reference's cosine pre-weight (pixel-index offsets, source row zOrigin),
    # ramp filter and backprojection with the default kernelMode 2 (Feldkamp) and norm 1/N, on
    # the whole detector; proj is filtered in place as in fastcone.py:260-270
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(47)
    nv, Nh, Nw = 9, 26, 30
    Nz, Ny, Nx = 16, 18, 20
    ang = np.sort(rng.uniform(0, 2 * np.pi, nv)).astype(np.float32)
    z0 = rng.uniform(Nz / 2 - 2, Nz / 2 + 2, nv).astype(np.float32)
    prob = oracle.Problem(ang.astype(np.float64), z0.astype(np.float64), 70.0, 0.8,
                          vol_dim=(Nx, Ny, Nz), proj_dim=(Nw, Nh))
    y = rng.standard_normal((nv, Nh, Nw)).astype(np.float32)
    proj = y.copy()
    vol = np.zeros((Nz, Ny, Nx), np.float32)
    cone.fbp(vol, proj, ang, 70.0, z0, pixelWidth=0.8)
    yf = oracle.rampfilter(y.astype(np.float64) * oracle.fbp_cosine_weight(y.shape, 70.0, z0))
    ref = oracle.backproject(prob, yf, vol.shape, mode=2, norm=np.full(nv, 1.0 / nv))
    assert rel(proj, yf) < 1e-5
    assert rel(vol, ref) < 5e-5


def test_abi_tiling_invariance(gpu):
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(46)
    c = _abi_case(rng)
    kw = dict(volStart=c["vs"], volDims=c["dims"], projWOffset=c["woff"], projHOffset=c["hoff"],
              projDims=c["pdims"], pixelWidth=0.8)
    p_whole = np.zeros((len(c["ang"]),) + c["win"], np.float32)
    cone.project(c["vol"], p_whole, c["ang"], 70.0, c["z0"], **kw)
    v_whole = np.zeros_like(c["vol"])
    cone.backproject(v_whole, p_whole, c["ang"], 70.0, c["z0"], **kw)
    try:
        cone.set_tile_limits(max_voxels=22 * 19 * 3, max_views=3)
        p_tiled = np.zeros_like(p_whole)
        cone.project(c["vol"], p_tiled, c["ang"], 70.0, c["z0"], **kw)
        v_tiled = np.zeros_like(c["vol"])
        cone.backproject(v_tiled, p_whole, c["ang"], 70.0, c["z0"], **kw)
    finally:
        cone.set_tile_limits(0, 0)
    assert rel(p_tiled, p_whole) < 1e-6
    assert rel(v_tiled, v_whole) < 1e-6


def test_abi_bad_device_raises(gpu):
    from fastcone_amd import fastcone as cone
    rng = np.random.default_rng(47)
    c = _abi_case(rng, nv=2)
    proj = np.zeros((2,) + c["win"], np.float32)
    with pytest.raises(_native.FastconeError) as e:
        cone.project(c["vol"], proj, c["ang"], 70.0, c["z0"], GPUID=999)
    assert e.value.code == _native.FC_ERR_DEVICE
    assert np.all(proj == 0)


# ---- configuration 0 at full size ----------------------------------------------------------------

def test_config0_full_size_parity(gpu):
    """128^3 phantom, 180 views over 360 deg, 192^2 detector, pitch 0.75 voxel at the axis,
    D = 500 voxels, step 0.5: FP, the device FDK filter and FDK backprojection each within 1e-4
    of the FP64 oracle, and the whole FDK chain (device filter + backprojection of the GPU's
    projections) against the oracle's FP64 filter + backprojection of the same projections."""
    from fastcone_amd import ops, phantom
    from fastcone_amd.geometry import Geometry
    vol, _ = phantom.config_phantom(0, device=gpu)
    ang = 2 * np.pi * np.arange(180) / 180
    g = Geometry.circular(ang, 500.0, 0.75, (128, 128, 128), (192, 192))
    prob = oracle.Problem(ang, np.full(180, 64.0), 500.0, 0.75, vol_dim=(128, 128, 128),
                          proj_dim=(192, 192))
    p = ops.project(vol, g, step=0.5)
    vol_np = vol.cpu().numpy()
    ref = oracle.project(prob, vol_np, h=0.5)
    e_fp = rel(p.cpu(), ref)
    q = ops.fdk_filter(p, g, 500.0, 0.75)
    q_ref = oracle.fdk_filter(prob, p.cpu().numpy())
    e_flt = rel(q.cpu(), q_ref)
    v = ops.backproject(q, g, mode=2, norm=torch.full((180,), 1 / 180, device=gpu)).cpu()
    ref_bp = oracle.backproject(prob, q.cpu().numpy(), (128, 128, 128), mode=2)
    e_bp = rel(v, ref_bp)
    e_chain = rel(v, oracle.backproject(prob, q_ref, (128, 128, 128), mode=2))
    print(f"config 0: FP rel L2 {e_fp:.2e}, filter {e_flt:.2e}, FDK BP {e_bp:.2e}, "
          f"filter + BP chain {e_chain:.2e}")
    assert e_fp < 1e-4 and e_flt < 5e-5 and e_bp < 1e-4 and e_chain < 1e-4
    # FDK reconstructs the phantom (inside the field of view) to the usual FDK accuracy
    inner = (slice(40, 88),) * 3
    err = np.linalg.norm(v.numpy()[inner] - vol_np[inner]) / np.linalg.norm(vol_np[inner])
    print(f"config 0: FDK vs phantom (central 48^3) rel L2 {err:.3f}")
    assert err < 0.35


# ---- empty and degenerate inputs -----------------------------------------------------------------

def test_empty_view_sets_are_no_ops(gpu):
    from fastcone_amd import ops
    from fastcone_amd.geometry import Geometry
    g = Geometry.circular(np.zeros(0), 60.0, 0.9, (8, 9, 10), (12, 14))
    vol = torch.ones(g.vol_shape, device=gpu)
    p = ops.project(vol, g, step=0.5)
    assert tuple(p.shape) == (0, 12, 14)
    out = torch.full(g.vol_shape, 2.0, device=gpu)
    ops.backproject(torch.zeros(g.proj_shape, device=gpu), g, mode=2, out=out)
    ops.adjoint(torch.zeros(g.proj_shape, device=gpu), g, step=0.5, out=out)
    torch.cuda.synchronize()
    assert torch.all(out == 2.0)


def test_zero_data_and_rays_missing_the_volume(gpu):
    # all-zero projections leave the adjoint's output untouched (its fixed-point scale comes from
    # max |y|); a detector window that sees none of the volume projects to exact zeros
    from fastcone_amd import ops
    rng = np.random.default_rng(48)
    g, prob = _setup(rng)
    out = torch.full(g.vol_shape, 3.0, device=gpu)
    ops.adjoint(torch.zeros(g.proj_shape, device=gpu), g, step=0.5, out=out)
    torch.cuda.synchronize()
    assert torch.all(out == 3.0)
    from fastcone_amd.geometry import Geometry
    far = Geometry.circular(np.array([0.3]), 60.0, 0.9, (8, 8, 8), (10, 10),
                            woff=np.array([380]), hoff=np.array([0]), det_dim=(10, 400))
    x = torch.rand(far.vol_shape, device=gpu)
    assert float(ops.project(x, far, step=0.5).abs().max()) == 0.0
    a = ops.adjoint(torch.ones(far.proj_shape, device=gpu), far, step=0.5)
    assert float(a.abs().max()) == 0.0


def test_abi_calls_with_no_views(gpu):
    from fastcone_amd import fastcone as cone
    vol = np.ones((6, 7, 8), np.float32)
    proj = np.zeros((0, 10, 12), np.float32)
    cone.project(vol, proj, np.zeros(0, np.float32), 70.0, np.zeros(0, np.float32))
    v = np.zeros_like(vol)
    cone.backproject(v, proj, np.zeros(0, np.float32), 70.0, np.zeros(0, np.float32),
                     norm=np.zeros(0, np.float32))
    assert np.all(v == 0)


@pytest.mark.parametrize("quad", [None, "always", "never"])
def test_project_random_dense_geometries(gpu, quad, monkeypatch):
    # seeded random geometries with adjacent views 0.1 - 2 deg apart (the quad layouts' domain):
    # windows, block offsets, steps, source distances down to 20 voxels and, in 40 % of the
    # cases, a tilted, shifted detector
    from fastcone_amd import ops
    from fastcone_amd.geometry import Geometry
    if quad is not None:
        monkeypatch.setenv("FASTCONE_FP_QUAD", quad)
    rng = np.random.default_rng(2024)
    for case in range(30):
        nv = int(rng.integers(1, 21))
        ang = rng.uniform(0, 2 * np.pi) + np.deg2rad(rng.uniform(0.1, 2.0)) * np.arange(nv)
        D, pw = float(rng.uniform(20.0, 150.0)), float(rng.uniform(0.5, 1.5))
        vol = tuple(int(v) for v in rng.integers(6, 40, 3))
        vol_dim = tuple(v + int(rng.integers(0, 4)) for v in vol)
        vs = tuple(int(rng.integers(0, d - v + 1)) for v, d in zip(vol, vol_dim))
        det = (int(rng.integers(6, 70)), int(rng.integers(6, 70)))
        det_dim = (det[0] + int(rng.integers(0, 6)), det[1] + int(rng.integers(0, 6)))
        hmid = rng.uniform(vol_dim[0] / 2 - 2, vol_dim[0] / 2 + 2, nv)
        woff, hoff = rng.integers(-3, 5, nv), rng.integers(-3, 5, nv)
        mis = None
        if rng.uniform() < 0.4:
            mis = list(rng.uniform(-1.0, 1.0, 3)) + list(rng.uniform(-0.3, 0.3, 3))
        h = float(rng.choice([0.37, 0.5, 1.0]))
        g = Geometry.circular(ang, D, pw, vol, det, hmid=hmid, vol_dim=vol_dim, vol_start=vs,
                              woff=woff, hoff=hoff, misalign=mis, det_dim=det_dim)
        prob = oracle.Problem(ang, hmid, D, pw, vol_dim=vol_dim[::-1], vol_start=vs[::-1],
                              proj_dim=det_dim[::-1], woff=woff, hoff=hoff, mis=mis)
        x = rng.standard_normal(vol).astype(np.float32)
        ref = oracle.project(prob, x, window=det[::-1], h=h)
        p = ops.project(torch.from_numpy(x).to(gpu), g, step=h).cpu()
        assert rel(p, ref) < 5e-5, (case, nv, D, pw, vol, det, mis, h)


def test_adjoint_of_non_finite_data_is_nan_and_cgls_raises(gpu):
    # a NaN or infinite datum on a ray that reaches the volume cannot be held by the adjoint's
    # fixed-point accumulator: the output becomes NaN (not silently zero), and CGLS stops with an
    # error instead of reporting convergence
    from fastcone_amd import cgls, ops
    rng = np.random.default_rng(49)
    g, prob = _setup(rng)
    base = torch.from_numpy(rng.standard_normal(g.proj_shape).astype(np.float32)).to(gpu)
    for bad in (float("nan"), float("inf")):
        y = base.clone()
        y[2, 13, 15] = bad   # a central pixel: its ray crosses the block
        out = ops.adjoint(y, g, step=0.5)
        torch.cuda.synchronize()
        assert bool(torch.isnan(out).all()), bad
    ok = ops.adjoint(base, g, step=0.5)
    assert bool(torch.isfinite(ok).all())
    y = base.clone()
    y[2, 13, 15] = float("nan")
    with pytest.raises(FloatingPointError):
        cgls.cgls(lambda x: ops.project(x, g, 0.5), lambda r: ops.adjoint(r, g, 0.5), y, iters=3)
