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
    # With no misalignment the general source-voxel-ray / detector-plane intersection of mode 0
    # (coneBeamGPU.cu:235-243) reduces to the FDK coordinates of coneBeamGPU.cu:398-399,
    # w = (y cos t - x sin t) U / pw, h = (z - zc) U / pw, U = D / (D + x cos t + y sin t), sampled
    # with no weight. Written here from the angles alone (no geometry vectors, no oracle code).
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(4)
    prob = _prob(rng)
    p = rng.standard_normal((7, 17, 19))
    shape = (9, 10, 12)
    a0 = oracle.backproject(prob, p, shape, mode=0)
    Nw, Nh = prob.proj_dim
    o = prob.origin
    zz, yy, xx = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    x, y, z = xx + o[0], yy + o[1], zz + o[2]
    ref = np.zeros(shape)
    for t, th in enumerate(prob.angles):
        U = prob.D / (prob.D + x * np.cos(th) + y * np.sin(th))
        w = (y * np.cos(th) - x * np.sin(th)) * U / prob.pw + Nw // 2 - prob.woff[t]
        h = (z - prob.zc[t]) * U / prob.pw + Nh // 2 - prob.hoff[t]
        ref += (1.0 / prob.n) * map_coordinates(p[t], np.stack([h.ravel(), w.ravel()]), order=1,
                                                mode="grid-constant", cval=0.0).reshape(shape)
    assert np.abs(a0 - ref).max() <= 1e-12 * np.abs(ref).max()
    # and the FDK mode is the same sampling weighted by U^2 (mode 1: U)
    a2 = oracle.backproject(prob, p, shape, mode=2)
    a1 = oracle.backproject(prob, p, shape, mode=1)
    ref2 = np.zeros(shape)
    ref1 = np.zeros(shape)
    for t, th in enumerate(prob.angles):
        U = prob.D / (prob.D + x * np.cos(th) + y * np.sin(th))
        w = (y * np.cos(th) - x * np.sin(th)) * U / prob.pw + Nw // 2 - prob.woff[t]
        h = (z - prob.zc[t]) * U / prob.pw + Nh // 2 - prob.hoff[t]
        b = map_coordinates(p[t], np.stack([h.ravel(), w.ravel()]), order=1,
                            mode="grid-constant", cval=0.0).reshape(shape)
        ref2 += (1.0 / prob.n) * U * U * b
        ref1 += (1.0 / prob.n) * U * b
    assert np.abs(a2 - ref2).max() <= 1e-12 * np.abs(ref2).max()
    assert np.abs(a1 - ref1).max() <= 1e-12 * np.abs(ref1).max()


# ---- the six-parameter misalignment chain (coneBeamGPU.cu:1701-1718) ------------------------------
# Each parameter alone is checked against what the mathematics fixes, with no oracle code: an
# integer detector shift is a window offset, Dl is a magnification, Dphi / Dtheta are rotations
# about the named axis (scipy's Rotation, in the sense the reference's cross products give), and
# Dpsi is the reference's in-plane "rotation", which the same identities show to be a skew.

def _geom(angles, D=40.0, pw=0.8, mis=None, hmid=6.0, nz=12):
    angles = np.asarray(angles, np.float64)
    return oracle.geometry(angles, np.full(angles.size, hmid), D, pw, nz, mis)


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_misalignment_aligned_vectors_closed_form():
    th = np.array([0.0, 0.4, 2.2, 5.0])
    src, det, ws, hs = _geom(th, D=40.0, pw=0.8, hmid=6.0, nz=12)
    zc = 6.0 - 12 // 2
    assert np.allclose(src, np.stack([-40 * np.cos(th), -40 * np.sin(th), np.full(4, zc)], 1))
    assert np.allclose(det, np.stack([0 * th, 0 * th, np.full(4, zc)], 1))
    assert np.allclose(ws, 0.8 * np.stack([-np.sin(th), np.cos(th), 0 * th], 1))
    assert np.allclose(hs, 0.8 * np.stack([0 * th, 0 * th, 1 + 0 * th], 1))


@pytest.mark.parametrize("axis,k", [(0, 3), (0, -2), (1, 2), (1, -4)])
def test_misalignment_integer_shift_equals_window_offset(axis, k):
    # Dw (Dh) of k pixels moves the detector centre by k w-steps (h-steps): window pixel j of the
    # shifted detector is pixel j + k of the unshifted one
    rng = np.random.default_rng(20 + axis)
    vol = rng.standard_normal((9, 10, 12))
    mis = np.zeros(6)
    mis[axis] = k
    a = _prob(rng, mis=mis)
    woff, hoff = a.woff.copy(), a.hoff.copy()
    if axis == 0:
        woff = woff + k
    else:
        hoff = hoff + k
    b = oracle.Problem(a.angles, a.hmid, a.D, a.pw, vol_dim=a.vol_dim, vol_start=a.vol_start,
                       proj_dim=a.proj_dim, woff=woff, hoff=hoff)
    pa = oracle.project(a, vol, window=(19, 17), h=0.5)
    pb = oracle.project(b, vol, window=(19, 17), h=0.5)
    assert np.abs(pa - pb).max() <= 1e-12 * np.abs(pb).max()


def test_misalignment_Dl_is_a_magnification():
    # the detector moved Dl further from the source along the central ray: every pixel ray equals
    # that of the aligned detector with pitch pw * D / (D + Dl)
    rng = np.random.default_rng(22)
    vol = rng.standard_normal((9, 10, 12))
    D, pw, Dl = 30.0, 0.8, 7.5
    a = _prob(rng, mis=[0, 0, Dl, 0, 0, 0], D=D, pw=pw)
    b = oracle.Problem(a.angles, a.hmid, D, pw * D / (D + Dl), vol_dim=a.vol_dim,
                       vol_start=a.vol_start, proj_dim=a.proj_dim, woff=a.woff, hoff=a.hoff)
    pa = oracle.project(a, vol, window=(19, 17), h=0.5)
    pb = oracle.project(b, vol, window=(19, 17), h=0.5)
    assert np.abs(pa - pb).max() <= 1e-11 * np.abs(pb).max()
    src, det, ws, hs = _geom(a.angles, D=D, pw=pw, mis=[0, 0, Dl, 0, 0, 0])
    assert np.allclose(np.linalg.norm(det - src, axis=1), D + Dl)


def test_misalignment_Dl_follows_the_shifted_centre():
    # Dl moves the (already Dw/Dh-shifted) centre along the source -> centre line
    th = np.array([0.3, 1.9])
    mis = [2.0, -1.0, 5.0, 0, 0, 0]
    src, det, ws, hs = _geom(th, mis=mis)
    src0, det0, ws0, hs0 = _geom(th)
    shifted = det0 + 2.0 * ws0 - 1.0 * hs0
    assert np.allclose(np.linalg.norm(det - src, axis=1), np.linalg.norm(shifted - src, axis=1) + 5.0)
    assert np.allclose(_unit(det - src), _unit(shifted - src))


def test_misalignment_Dphi_rotates_h_about_w():
    from scipy.spatial.transform import Rotation
    th = np.array([0.0, 0.7, 3.5])
    phi = 0.21
    src0, det0, ws0, hs0 = _geom(th)
    src, det, ws, hs = _geom(th, mis=[0, 0, 0, phi, 0, 0])
    for t in range(th.size):
        R = Rotation.from_rotvec(phi * _unit(ws0[t]))
        assert np.allclose(hs[t], R.apply(hs0[t]), atol=1e-12)
        assert np.allclose(ws[t], ws0[t], atol=1e-12)
    assert np.allclose(src, src0) and np.allclose(det, det0)
    # positive Dphi tilts h away from the source
    assert np.all(np.einsum("ij,ij->i", hs, det0 - src0) > 0)


def test_misalignment_Dtheta_rotates_w_about_h():
    from scipy.spatial.transform import Rotation
    th = np.array([0.0, 0.7, 3.5])
    tt = -0.17
    src0, det0, ws0, hs0 = _geom(th)
    src, det, ws, hs = _geom(th, mis=[0, 0, 0, 0, tt, 0])
    for t in range(th.size):
        # w <- cos(tt) w + sin(tt) (w x h)^: a rotation about h by -tt
        R = Rotation.from_rotvec(-tt * _unit(hs0[t]))
        assert np.allclose(ws[t], R.apply(ws0[t]), atol=1e-12)
        assert np.allclose(hs[t], hs0[t], atol=1e-12)


def test_misalignment_Dpsi_is_a_skew_of_w_and_h():
    # w <- cos(psi) w - sin(psi) (w x n)^ and h <- cos(psi) h + sin(psi) (h x n)^, n = w x h:
    # for orthonormal w, h the identities a x (b x c) = b (a.c) - c (a.b) give (w x n)^ = -h and
    # (h x n)^ = w, so both vectors tilt towards each other (not a rotation: w'.h' = sin 2 psi)
    th = np.array([0.2, 2.6])
    psi = 0.13
    src0, det0, ws0, hs0 = _geom(th, pw=0.8)
    src, det, ws, hs = _geom(th, pw=0.8, mis=[0, 0, 0, 0, 0, psi])
    w, h = _unit(ws0), _unit(hs0)
    assert np.allclose(ws, 0.8 * (np.cos(psi) * w + np.sin(psi) * h), atol=1e-12)
    assert np.allclose(hs, 0.8 * (np.sin(psi) * w + np.cos(psi) * h), atol=1e-12)
    assert np.allclose(np.einsum("ij,ij->i", _unit(ws), _unit(hs)), np.sin(2 * psi))


def test_misalignment_full_chain_composes_in_order():
    # Dw, Dh, Dl on the centre; then Dphi (h about w), Dtheta (w about the new h), Dpsi (skew of
    # the resulting pair): composed with scipy rotations, against the oracle's chain
    from scipy.spatial.transform import Rotation
    th = np.array([0.5, 4.1])
    mis = np.array([0.6, -0.8, 1.5, 0.09, -0.12, 0.07])
    src0, det0, ws0, hs0 = _geom(th, D=50.0, pw=0.9)
    src, det, ws, hs = _geom(th, D=50.0, pw=0.9, mis=mis)
    for t in range(th.size):
        d = det0[t] + mis[0] * ws0[t] + mis[1] * hs0[t]
        d = d + mis[2] * _unit(d - src0[t])
        h1 = Rotation.from_rotvec(mis[3] * _unit(ws0[t])).apply(hs0[t])
        w1 = Rotation.from_rotvec(-mis[4] * _unit(h1)).apply(ws0[t])
        w2 = 0.9 * (np.cos(mis[5]) * _unit(w1) + np.sin(mis[5]) * _unit(h1))
        h2 = 0.9 * (np.sin(mis[5]) * _unit(w1) + np.cos(mis[5]) * _unit(h1))
        assert np.allclose(src[t], src0[t], atol=1e-12)
        assert np.allclose(det[t], d, atol=1e-12)
        assert np.allclose(ws[t], w2, atol=1e-12)
        assert np.allclose(hs[t], h2, atol=1e-12)
    # pitch lengths are preserved by every step
    assert np.allclose(np.linalg.norm(ws, axis=1), 0.9)
    assert np.allclose(np.linalg.norm(hs, axis=1), 0.9)


# ---- physical-units FDK pre-processing ------------------------------------------------------------

def test_fdk_cosine_weight_closed_form_when_aligned():
    # aligned detector through the axis: the weight is D / sqrt(D^2 + (gw pw)^2 + (gh pw)^2) with
    # the global pixel offsets gw = j + woff - Nw//2, gh = i + hoff - Nh//2
    rng = np.random.default_rng(30)
    prob = _prob(rng, D=30.0, pw=0.8)
    w = oracle.fdk_cosine_weight(prob, window=(19, 17))
    Nw, Nh = prob.proj_dim
    for t in range(prob.n):
        gw = np.arange(19) + prob.woff[t] - Nw // 2
        gh = np.arange(17) + prob.hoff[t] - Nh // 2
        ref = 30.0 / np.sqrt(30.0 ** 2 + (0.8 * gw[None, :]) ** 2 + (0.8 * gh[:, None]) ** 2)
        assert np.allclose(w[t], ref, rtol=0, atol=1e-14)


def test_fdk_cosine_weight_tilted_detector_is_the_ray_normal_cosine():
    # tilted detector: w |P - src| is the same for every pixel (the source's distance from the
    # detector plane, here from a least-squares fit of the plane through the pixel centres)
    rng = np.random.default_rng(31)
    prob = _prob(rng, mis=[0.7, -1.2, 2.0, 0.2, -0.25, 0.0])
    w = oracle.fdk_cosine_weight(prob, window=(19, 17))
    Nw, Nh = prob.proj_dim
    for t in range(prob.n):
        jj, ii = np.meshgrid(np.arange(19) + prob.woff[t] - Nw // 2,
                             np.arange(17) + prob.hoff[t] - Nh // 2)
        P = prob.det[t] + jj[..., None] * prob.wstep[t] + ii[..., None] * prob.hstep[t]
        Q = P.reshape(-1, 3)
        c = Q.mean(0)
        nrm = np.linalg.svd(Q - c)[2][-1]
        dist = abs(np.dot(c - prob.src[t], nrm))
        assert np.allclose(w[t] * np.linalg.norm(P - prob.src[t], axis=-1), dist, rtol=1e-12)
        assert w[t].max() <= 1.0


def test_fdk_filter_is_weight_then_ramp_over_pw():
    # linear in the data; an impulse in row i, column m gives weight(i, m) / pw times the ramp's
    # impulse response shifted to m (the explicit inverse DFT of the 2 Nw ramp)
    rng = np.random.default_rng(32)
    prob = _prob(rng, n=2, D=30.0, pw=0.8)
    nw = 19
    p = np.zeros((2, 17, nw))
    p[1, 5, 7] = 1.0
    q = oracle.fdk_filter(prob, p, window=(nw, 17))
    w = oracle.fdk_cosine_weight(prob, window=(nw, 17))
    n = 2 * nw
    H = oracle.ramp_filter_response(nw)
    Hf = np.concatenate([H, H[1:-1][::-1]])
    k = np.arange(n)
    h = np.array([np.sum(Hf * np.cos(2 * np.pi * k * d / n)) / n for d in range(n)])
    ref = w[1, 5, 7] / 0.8 * h[(np.arange(nw) - 7) % n]
    assert np.allclose(q[1, 5], ref, atol=1e-14)
    assert np.abs(q[0]).max() == 0.0 and np.abs(np.delete(q[1], 5, axis=0)).max() == 0.0


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
