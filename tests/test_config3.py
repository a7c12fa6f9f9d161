"""Configuration 3: the differentiable projector pair at full size (256^3, 360 views over 360 deg,
384^2 detector, pitch 1.2 mm = 0.6 voxel at the axis, D = 500 voxels, step 0.5 voxel).

* dot-product test <Ax, y> = <x, A^T y> on seeded N(0,1) x and y, FP64 dots;
* central finite differences (h = 1e-3) of L(x) = <A x, y> on 64 random voxels against the
  autograd gradient (the matched adjoint);
* 50 CGLS iterations on Poisson-noised data (I0 = 1e4 photons per ray, log transform) with FP64
  inner products: |r_k| non-increasing, normal-equation residual reduced, reconstruction error
  reported.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N, NV, ND, D, PW, H = 256, 360, 384, 500.0, 0.6, 0.5


@pytest.fixture(scope="module")
def geom():
    from fastcone_amd.geometry import Geometry
    ang = 2 * np.pi * np.arange(NV) / NV
    return Geometry.circular(ang, D, PW, (N, N, N), (ND, ND))


def test_dot_product_and_finite_differences(gpu, geom):
    from fastcone_amd import ops
    gen = torch.Generator(device=gpu)
    gen.manual_seed(3)
    x = torch.randn((N, N, N), generator=gen, device=gpu)
    y = torch.randn(geom.proj_shape, generator=gen, device=gpu)
    xr = x.clone().requires_grad_(True)
    ax = ops.cone_project(xr, geom, H)
    loss = (ax.double() * y.double()).sum()
    loss.backward()
    aty = xr.grad
    lhs = float(loss)
    rhs = float((x.double() * aty.double()).sum())
    gap = abs(lhs - rhs) / abs(lhs)
    print(f"config 3 dot-product test: <Ax,y> = {lhs:.9e}, <x,A^T y> = {rhs:.9e}, gap {gap:.2e}")
    assert gap < 1e-5
    # central differences on 64 seeded voxels
    rng = np.random.default_rng(4)
    hh = 1e-3
    idx = [tuple(int(v) for v in rng.integers(N // 4, 3 * N // 4, 3)) for _ in range(64)]
    fd, an = [], []
    with torch.no_grad():
        for (z, yy, xx) in idx:
            xp = x.clone()
            xp[z, yy, xx] += hh
            lp = float((ops.project(xp, geom, H).double() * y.double()).sum())
            xp[z, yy, xx] -= 2 * hh
            lm = float((ops.project(xp, geom, H).double() * y.double()).sum())
            fd.append((lp - lm) / (2 * hh))
            an.append(float(aty[z, yy, xx]))
    fd, an = np.array(fd), np.array(an)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    print(f"config 3 finite differences (h=1e-3, 64 voxels): rel L2 {err:.2e}")
    assert err < 5e-2


def test_cgls_50_iterations_poisson(gpu, geom):
    from fastcone_amd import ops, phantom
    from fastcone_amd.cgls import cgls, poisson_log
    vol, _ = phantom.config_phantom(3, device=gpu)
    p = ops.project(vol, geom, H)
    b = poisson_log(p, 1e4, scale=1.0, seed=5)  # voxel = 1 mm: line integrals are in mm
    A = lambda v: ops.project(v, geom, H)
    At = lambda r: ops.adjoint(r, geom, H)
    torch.cuda.synchronize()
    x, info = cgls(A, At, b, iters=50)
    torch.cuda.synchronize()
    rn = np.array(info["rnorm"])
    res = np.array(info["residual"])
    err = float(torch.linalg.norm(x - vol) / torch.linalg.norm(vol))
    print(f"config 3 CGLS: 50 iterations in {info['time_s']:.2f} s, |r| {rn[0]:.3e} -> {rn[-1]:.3e},"
          f" normal residual {res[-1]:.3e}, rel error to phantom {err:.3f}")
    assert info["iters"] == 50
    assert np.all(np.diff(rn) <= 1e-5 * rn[0])
    assert res[-1] < 0.05
    assert err < 0.5
