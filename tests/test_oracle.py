"""Pins for the FP64 oracle (CPU). Each test checks the oracle against something it does not
share code with: a library routine (scipy map_coordinates), a closed form, an exact invariant
(adjointness, rotation symmetry), or a brute-force enumeration."""
import math

import numpy as np
import pytest

import oracle


def _prob(rng, n=7, vol_dim=(16, 14, 11), vol_start=(2, 1, 1), proj_dim=(20, 18), mis=None,
          D=30.0, pw=0.8):
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False) + 0.1
    return oracle.Problem(ang, hmid=np.full(n, 5.3), D=D, pw=pw, vol_dim=vol_dim,
                          vol_start=vol_start, proj_dim=proj_dim,
                          woff=rng.integers(-2, 3, n), hoff=rng.integers(-2, 3, n), mis=mis)


@pytest.mark.parametrize("h", [1.0, 0.5, 0.37])
def test_project_matches_scipy_bruteforce(h):
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((9, 10, 12))
    prob = _prob(rng)
    a = oracle.project(prob, vol, window=(19, 17), h=h)
    b = oracle.project_bruteforce(prob, vol, window=(19, 17), h=h)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_project_misaligned_matches_bruteforce():
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((9, 10, 12))
    prob = _prob(rng, mis=[0.7, -1.2, 2.0, 0.05, -0.08, 0.11])
    a = oracle.project(prob, vol, window=(19, 17), h=0.5)
    b = oracle.project_bruteforce(prob, vol, window=(19, 17), h=0.5)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_backproject_matches_projective_bruteforce(mode):
    rng = np.random.default_rng(2)
    prob = _prob(rng)
    p = rng.standard_normal((7, 17, 19))
    a = oracle.backproject(prob, p, (9, 10, 12), mode=mode)
    b = oracle.backproject_bruteforce(prob, p, (9, 10, 12), mode=mode)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_backproject_mode0_misaligned_matches_bruteforce():
    rng = np.random.default_rng(3)
    prob = _prob(rng, mis=[0.7, -1.2, 2.0, 0.05, -0.08, 0.11])
    p = rng.standard_normal((7, 17, 19))
    a = oracle.backproject(prob, p, (9, 10, 12), mode=0)
    b = oracle.backproject_bruteforce(prob, p, (9, 10, 12), mode=0)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_mode0_equals_unweighted_fdk_geometry_when_aligned():
    # with no misalignment the general intersection reduces to the FDK coordinates
    rng = np.random.default_rng(4)
    prob = _prob(rng)
    p = rng.standard_normal((7, 17, 19))
    a0 = oracle.backproject(prob, p, (9, 10, 12), mode=0)
    b0 = oracle.backproject_bruteforce(prob, p, (9, 10, 12), mode=0)
    assert np.allclose(a0, b0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("h", [1.0, 0.5])
def test_adjoint_dot_product(h):
    rng = np.random.default_rng(5)
    prob = _prob(rng, mis=[0.3, 0.2, -0.5, 0.02, 0.01, -0.03])
    x = rng.standard_normal((9, 10, 12))
    y = rng.standard_normal((7, 17, 19))
    norm = rng.uniform(0.5, 1.5, 7)
    Ax = oracle.project(prob, x, window=(19, 17), h=h, norm=norm)
    ATy = oracle.adjoint(prob, y, x.shape, h=h, norm=norm)
    lhs, rhs = np.vdot(Ax, y), np.vdot(x, ATy)
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(Ax).sum() * np.abs(y).max())


@pytest.mark.parametrize("h", [1.0, 0.5])
def test_central_ray_is_exact_row_sum(h):
    # Source at (-D, 0, 0) (theta = 0, hMid = Nz/2): the central pixel's ray runs along +x through
    # voxel centres y = z = 0. With D a multiple of h the samples hit integer and (h = 0.5)
    # half-integer x, where the zero-extended trilinear interpolant is v[i] or a neighbour mean, so
    # h * sum_k T = sum_i v[i] exactly.
    rng = np.random.default_rng(6)
    N = 16
    vol = rng.standard_normal((N, N, N))
    prob = oracle.Problem([0.0], hmid=[N // 2], D=40.0, pw=1.0, vol_dim=(N, N, N),
                          proj_dim=(N, N))
    p = oracle.project(prob, vol, h=h)
    row = vol[N // 2, N // 2, :].sum()
    assert abs(p[0, N // 2, N // 2] - row) <= 1e-12 * np.abs(vol).sum()


def test_constant_volume_chord_lengths():
    # Constant 1 on the block: the zero-extended trilinear interpolant of an axis-aligned ray through
    # voxel centres integrates to exactly N (tent ramps of area 1/2 at both ends).
    N = 12
    vol = np.ones((N, N, N))
    prob = oracle.Problem([0.0, np.pi / 2], hmid=[N // 2] * 2, D=60.0, pw=1.0,
                          vol_dim=(N, N, N), proj_dim=(N, N))
    p = oracle.project(prob, vol, h=0.25)
    assert abs(p[0, N // 2, N // 2] - N) < 1e-9
    assert abs(p[1, N // 2, N // 2] - N) < 1e-9


def test_rotation_by_90_degrees():
    # Odd N keeps the coordinate grid symmetric: rotating the volume by +90 deg about z and the
    # view angle by +90 deg leaves every projection unchanged.
    rng = np.random.default_rng(7)
    N = 11
    vol = rng.standard_normal((7, N, N))
    th = np.array([0.3, 1.1, 2.0])
    kw = dict(D=35.0, pw=0.9, vol_dim=(N, N, 7), proj_dim=(15, 13))
    p1 = oracle.project(oracle.Problem(th, hmid=[3.4] * 3, **kw), vol, h=0.5)
    # vol_rot(x, y) = vol(y, -x): vol_rot[:, Y, X] = vol[:, N-1-X, Y]
    Y, X = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    vol_rot = vol[:, N - 1 - X, Y]
    p2 = oracle.project(oracle.Problem(th + np.pi / 2, hmid=[3.4] * 3, **kw), vol_rot, h=0.5)
    assert np.abs(p1 - p2).max() < 1e-9 * np.abs(p1).max()


def test_linearity():
    rng = np.random.default_rng(8)
    prob = _prob(rng)
    a, b = rng.standard_normal((2, 9, 10, 12))
    pa = oracle.project(prob, a, window=(19, 17), h=0.5)
    pb = oracle.project(prob, b, window=(19, 17), h=0.5)
    pab = oracle.project(prob, 2.0 * a - 3.0 * b, window=(19, 17), h=0.5)
    assert np.allclose(pab, 2 * pa - 3 * pb, atol=1e-10)


def test_ellipsoid_line_integrals_closed_form():
    # A large smooth-ish ellipsoid: the voxelised projection approaches the analytic chord
    # integral (voxelisation error only; a dropped h factor or wrong pitch fails by >> 2%).
    from fastcone_amd import phantom
    N = 48
    R = phantom.random_rotation(np.random.default_rng(9))
    ell = [(np.array([1.5, -2.0, 0.5]), np.array([15.0, 11.0, 9.0]), R, 1.0)]
    vol = phantom.voxelize(ell, (N, N, N), 1.0).numpy().astype(np.float64)
    ang = np.array([0.2, 1.3])
    prob = oracle.Problem(ang, hmid=[N // 2] * 2, D=100.0, pw=1.0, vol_dim=(N, N, N),
                          proj_dim=(N, N))
    p = oracle.project(prob, vol, h=0.25)
    for t in range(2):
        js, is_ = np.meshgrid(np.arange(N), np.arange(N))
        P = (prob.det[t] + (js[..., None] - N // 2) * prob.wstep[t]
             + (is_[..., None] - N // 2) * prob.hstep[t])
        ana = oracle.ellipsoid_line_integrals(prob.src[t], P, ell)
        mask = ana > 0.5 * ana.max()
        rel = np.abs(p[t][mask] - ana[mask]) / ana[mask]
        assert np.median(rel) < 0.02
        assert np.linalg.norm(p[t] - ana) / np.linalg.norm(ana) < 0.05


def test_sample_count_bruteforce():
    rng = np.random.default_rng(10)
    prob = _prob(rng, n=3)
    shape = (9, 10, 12)
    h = 0.5
    counts = oracle.sample_count(prob, shape, window=(19, 17), h=h)
    nz, ny, nx = shape
    o = prob.origin
    for t in range(3):
        c = 0
        for i in range(17):
            for j in range(19):
                P = (prob.det[t] + (j + prob.woff[t] - 10) * prob.wstep[t]
                     + (i + prob.hoff[t] - 9) * prob.hstep[t])
                u = P - prob.src[t]
                u /= np.linalg.norm(u)
                ks = np.arange(-400, 400) * h
                q = (prob.src[t] - o)[None, :] + ks[:, None] * u[None, :]
                inside = np.all((q >= -1) & (q <= np.array([nx, ny, nz])), axis=1)
                c += int(inside.sum())
        assert c == counts[t]


# ---- FDK pre-processing -----------------------------------------------------------------------

def test_ramp_impulse_response_is_the_inverse_dft():
    # Filtering a unit impulse at column m0 gives h[(j - m0) mod n] cropped to [0, Nw), where
    # h = inverse DFT of the real, even ramp: h[m] = (1/n)(H0 + 2 sum_{k=1}^{n/2-1} H_k cos(2 pi k m / n)
    # + H_{n/2} cos(pi m)) -- evaluated here by explicit cosine sums, not by an FFT.
    nw = 13
    n = 2 * nw
    H = oracle.ramp_filter_response(nw)
    m = np.arange(n)
    k = np.arange(1, n // 2)
    h = (H[0] + 2 * (H[k][None, :] * np.cos(2 * np.pi * np.outer(m, k) / n)).sum(1)
         + H[n // 2] * np.cos(np.pi * m)) / n
    for m0 in (0, 5, nw - 1):
        x = np.zeros((1, 1, nw))
        x[0, 0, m0] = 1.0
        y = oracle.rampfilter(x)[0, 0]
        expect = h[(np.arange(nw) - m0) % n]
        assert np.allclose(y, expect, atol=1e-14)


def test_ramp_dc_and_values():
    nw = 10
    H = oracle.ramp_filter_response(nw)
    n = 2 * nw
    assert H.size == nw + 1
    assert abs(H[0] - np.pi / (16 * nw)) < 1e-15
    assert np.allclose(H[1:], np.pi * np.arange(1, nw + 1) / n)


def test_fbp_cosine_weight_closed_form():
    w = oracle.fbp_cosine_weight((2, 5, 8), 10.0, [2.0, 3.0])
    # centre column (j = Nw/2) and row i = zOrigin: weight 1
    assert abs(w[0, 2, 4] - 1.0) < 1e-15 and abs(w[1, 3, 4] - 1.0) < 1e-15
    assert abs(w[0, 0, 0] - 10.0 / math.sqrt(100 + 4 + 16)) < 1e-15


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_backproject_constant_projection_closed_form(mode):
    # constant-1 data on a detector that every voxel projects well inside: the bilinear sample is
    # 1, so each voxel receives the sum over views of its weight: U^2 (FDK), U (Katsevich) or 1,
    # with U = D / (D + x cos t + y sin t) and x, y measured from the volume centre (index N // 2)
    ang = np.linspace(0.3, 0.3 + 2 * np.pi, 7, endpoint=False)
    D, nz, ny, nx = 40.0, 6, 9, 11
    prob = oracle.Problem(ang, np.full(7, nz / 2.0), D, 1.0, vol_dim=(nx, ny, nz),
                          proj_dim=(64, 64))
    v = oracle.backproject(prob, np.ones((7, 64, 64)), (nz, ny, nx), mode=mode, norm=np.ones(7))
    x = np.arange(nx) - nx // 2
    y = np.arange(ny) - ny // 2
    U = D / (D + x[None, None, :] * np.cos(ang)[:, None, None]
             + y[None, :, None] * np.sin(ang)[:, None, None])
    expect = {0: np.full(U.shape, 1.0), 1: U, 2: U ** 2}[mode].sum(0)
    assert np.allclose(v, np.broadcast_to(expect, v.shape), rtol=1e-12, atol=1e-12)


def test_backproject_single_pixel_delta():
    # one view, a unit delta at detector pixel (w0, h0): voxel (x, y, z) receives U^2 times the
    # bilinear hat weight (1 - |w - w0|)(1 - |h - h0|) of its projected position
    ang = np.array([0.37])
    D, nz, ny, nx, nw, nh = 30.0, 8, 10, 12, 24, 20
    prob = oracle.Problem(ang, np.array([nz / 2.0]), D, 1.0, vol_dim=(nx, ny, nz),
                          proj_dim=(nw, nh))
    w0, h0 = 13, 9
    p = np.zeros((1, nh, nw))
    p[0, h0, w0] = 1.0
    v = oracle.backproject(prob, p, (nz, ny, nx), mode=2, norm=np.ones(1))
    z, y, x = np.meshgrid(np.arange(nz) - nz // 2, np.arange(ny) - ny // 2,
                          np.arange(nx) - nx // 2, indexing="ij")
    c, s = math.cos(ang[0]), math.sin(ang[0])
    U = D / (D + x * c + y * s)
    w = (y * c - x * s) * U + nw // 2
    h = (z - (nz / 2.0 - nz // 2)) * U + nh // 2
    hat = np.clip(1 - np.abs(w - w0), 0, None) * np.clip(1 - np.abs(h - h0), 0, None)
    assert np.count_nonzero(hat) > 4
    assert np.allclose(v, U ** 2 * hat, rtol=1e-10, atol=1e-12)
