"""Helical cone beam on a cylindrical detector and exact Katsevich reconstruction.

CPU: the oracle's helical projector against closed-form ellipsoid line integrals, and the
oracle's Katsevich reconstruction of analytic projections against the phantom (exactness of the
method inside the reconstructable region pins every filtering stage, its sign and its scale).
GPU: libFastcone's helical projector and fc_katsevich_dev against the oracle on the same data,
then configuration 2 at full size.
"""
import math

import numpy as np
import pytest

import oracle

# a small helix: R = 100, SDD = 200 voxels, 96 x 32 curved detector, pitch 0.8 x detector height
R, SDD, NW, NH, DG, DROW = 100.0, 200.0, 96, 32, 0.0052, 1.0
PITCH = 0.8 * NH * DROW * R / SDD
HP = PITCH / (2 * math.pi)
VPT, TURNS = 360, 3
N = 32
ORIGIN = np.array([-16.0, -16.0, -16.0])
ELL = [(np.array([1.5, -2.0, 0.5]), np.array([12.0, 10.0, 9.0]), np.eye(3), 1.0),
       (np.array([-4.0, 3.0, 1.0]), np.array([3.0, 3.0, 3.0]), np.eye(3), 0.5)]


def small_helix():
    lam = 2 * math.pi * np.arange(VPT * TURNS) / VPT
    zs = HP * (lam - lam.mean())
    return oracle.Helix(lam, zs, R, SDD, DG, DROW, NW, NH, HP)


def analytic_projections(hx):
    proj = np.zeros((hx.n, hx.nh, hx.nw))
    for t in range(hx.n):
        a, pts = hx.ray_points(t)
        proj[t] = oracle.ellipsoid_line_integrals(a, pts, ELL)
    return proj


def phantom_grid():
    zz, yy, xx = np.meshgrid(*(np.arange(N) + o for o in ORIGIN[::-1]), indexing="ij")
    truth = np.zeros((N, N, N))
    inner = np.zeros((N, N, N), bool)
    for c, ax, _, mu in ELL:
        q = ((xx - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2 + ((zz - c[2]) / ax[2]) ** 2
        truth += mu * (q <= 1)
    # voxels at least ~2 voxels from every ellipsoid surface, inside the reconstructable slab
    d = np.full((N, N, N), np.inf)
    for c, ax, _, mu in ELL:
        q = np.sqrt(((xx - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2
                    + ((zz - c[2]) / ax[2]) ** 2)
        d = np.minimum(d, np.abs(q - 1) * ax.min())
    inner = (d > 2.0) & (np.abs(zz) < 6) & (np.hypot(xx, yy) < 14)
    return truth, inner


def test_oracle_helical_projector_matches_analytic():
    from fastcone_amd import phantom
    hx = small_helix()
    sub = oracle.Helix(hx.lambdas[::90], hx.zs[::90], R, SDD, DG, DROW, NW, NH, HP)
    ana = analytic_projections(sub)
    vol = phantom.voxelize(ELL, (N, N, N), 1.0).numpy().astype(np.float64)
    p = oracle.project_helical(sub, vol, ORIGIN, h=0.25)
    mask = ana > 0.5 * ana.max()
    assert np.median(np.abs(p[mask] - ana[mask]) / ana[mask]) < 0.02
    assert np.linalg.norm(p - ana) / np.linalg.norm(ana) < 0.06


def test_oracle_katsevich_reconstructs_analytic_phantom():
    hx = small_helix()
    proj = analytic_projections(hx)
    rec = oracle.katsevich(hx, proj, (N, N, N), ORIGIN)
    truth, inner = phantom_grid()
    err = np.abs(rec - truth)[inner]
    print(f"oracle Katsevich: interior max |err| {err.max():.3f}, mean {err.mean():.4f}")
    assert inner.sum() > 2000
    assert err.mean() < 0.01 and err.max() < 0.08


# ---- GPU -------------------------------------------------------------------------------------

def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_helical_projector_parity(gpu):
    import torch
    from fastcone_amd import ops
    from fastcone_amd.katsevich import HelixScan
    hx = small_helix()
    scan = HelixScan(hx.n, hx.dl, hx.lambdas[0], hx.zs[0], HP, R, SDD, DG, DROW, NW, NH)
    g = scan.geometry((N, N, N))
    sub = slice(0, hx.n, 37)
    gs = g.subset(sub)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((N, N, N)).astype(np.float32)
    p = ops.project(torch.from_numpy(x).to(gpu), gs, step=0.5).cpu().numpy()
    hs = oracle.Helix(hx.lambdas[sub], hx.zs[sub], R, SDD, DG, DROW, NW, NH, HP)
    ref = oracle.project_helical(hs, x, ORIGIN, h=0.5)
    assert _rel(p, ref) < 5e-5



@pytest.mark.gpu
def test_helical_adjoint_dot_product(gpu):
    # cylindrical records take the adjoint's whole-detector windows; the pair must stay matched
    import torch
    from fastcone_amd import ops
    from fastcone_amd.katsevich import HelixScan
    hx = small_helix()
    scan = HelixScan(hx.n, hx.dl, hx.lambdas[0], hx.zs[0], HP, R, SDD, DG, DROW, NW, NH)
    gs = scan.geometry((N, N, N)).subset(slice(0, hx.n, 53))
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal((N, N, N)).astype(np.float32)).to(gpu)
    y = torch.from_numpy(rng.standard_normal(gs.proj_shape).astype(np.float32)).to(gpu)
    ax = ops.project(x, gs, step=0.5)
    aty = ops.adjoint(y, gs, step=0.5)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5

@pytest.mark.gpu
def test_katsevich_parity(gpu):
    import torch
    from fastcone_amd.katsevich import HelixScan, katsevich
    hx = small_helix()
    scan = HelixScan(hx.n, hx.dl, hx.lambdas[0], hx.zs[0], HP, R, SDD, DG, DROW, NW, NH)
    proj = analytic_projections(hx)
    rec, g5 = katsevich(torch.from_numpy(proj.astype(np.float32)).to(gpu), scan, (N, N, N),
                        origin=tuple(ORIGIN), return_filtered=True)
    ref, g5_ref = oracle.katsevich(hx, proj, (N, N, N), ORIGIN, return_g5=True)
    e_f = _rel(g5.cpu().numpy(), g5_ref)
    e_v = _rel(rec.cpu().numpy(), ref)
    truth, inner = phantom_grid()
    e_p = float(np.abs(rec.cpu().numpy() - truth)[inner].mean())
    print(f"Katsevich GPU vs FP64 oracle: filtered rel L2 {e_f:.2e}, volume rel L2 {e_v:.2e}; "
          f"interior mean |err| vs phantom {e_p:.4f}")
    assert e_f < 1e-4
    assert e_v < 1e-3
    assert e_p < 0.01
