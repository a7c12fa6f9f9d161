"""ORACLE — TEST INFRASTRUCTURE ONLY.

FP64 CPU restatement of the reference's cone-beam projector path. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this package, and
only as the checker (or the timed CPU baseline), never as the product path.

Contents
  * ``liboracle.so`` (oracle.cpp, C++/OpenMP FP64): geometry, forward projection at step h,
    FDK/Katsevich/unweighted backprojection and the exact transpose of the projector;
  * numpy FP64 restatements of ``fastcone.rampfilter``, of the ``fastcone.fbp`` cosine weight and
    of the physical-units FDK pre-processing (cosine weight from the geometry vectors + ramp / pw);
  * a scipy ``map_coordinates`` brute force for tiny problems (a second, independent
    implementation of the same interpolants: order-1 splines with ``mode="grid-constant"`` are the
    zero-extended trilinear / bilinear interpolants);
  * closed-form ellipsoid line integrals.

Every function is pinned by tests/test_oracle.py (closed forms, library routines, adjointness,
symmetry). There are no reference-held golden vectors for this path: the reference ships no tests
or fixtures, and its CUDA cannot be built here (no nvcc, no cutil_math.h, legacy texture API).
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile liboracle.so with the Makefile in this directory."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_num_threads.restype = ctypes.c_int
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def num_threads():
    return lib().orc_num_threads()


# ----------------------------------------------------------------------------------------------
# Geometry (coneBeamGPU.cu:1672-1734)
# ----------------------------------------------------------------------------------------------
def geometry(angles, hmid, D, pw, nz_total, mis=None):
    """Per-view (src, det, wstep, hstep), each [nviews, 3] float64, in voxel coordinates."""
    angles, pa = _d(angles)
    hmid, ph = _d(hmid)
    n = angles.size
    mis, pm = _d(np.zeros(6) if mis is None else mis)
    out = [np.zeros((n, 3)) for _ in range(4)]
    lib().orc_geometry(ctypes.c_int(n), pa, ph, ctypes.c_double(D), ctypes.c_double(pw), pm,
                       ctypes.c_int(int(nz_total)), *[o.ctypes.data_as(_dp) for o in out])
    return tuple(out)


class Problem:
    """A reference-convention problem description in C order [x,y,z] / [w,h,t].

    vol_dim   : whole-volume size [Nx, Ny, Nz] (fixes the centre at vol_dim // 2)
    vol_start : global index of the local block's voxel (0,0,0)
    proj_dim  : whole-detector size [Nw, Nh]
    woff/hoff : per-view global index of each window's pixel (0,0)
    """

    def __init__(self, angles, hmid, D, pw, vol_dim, vol_start=(0, 0, 0), proj_dim=None,
                 woff=None, hoff=None, mis=None):
        self.angles = np.asarray(angles, dtype=np.float64)
        self.n = self.angles.size
        self.hmid = np.broadcast_to(np.asarray(hmid, dtype=np.float64), (self.n,)).copy()
        self.D = float(D)
        self.pw = float(pw)
        self.vol_dim = np.asarray(vol_dim, dtype=np.int64)
        self.vol_start = np.asarray(vol_start, dtype=np.int64)
        self.proj_dim = np.asarray(proj_dim, dtype=np.int64)
        self.woff = np.zeros(self.n, np.int32) if woff is None else np.asarray(woff, np.int32)
        self.hoff = np.zeros(self.n, np.int32) if hoff is None else np.asarray(hoff, np.int32)
        self.mis = np.zeros(6) if mis is None else np.asarray(mis, dtype=np.float64)
        self.src, self.det, self.wstep, self.hstep = geometry(
            self.angles, self.hmid, self.D, self.pw, self.vol_dim[2], self.mis)

    @property
    def origin(self):
        return (self.vol_start - self.vol_dim // 2).astype(np.float64)

    @property
    def zc(self):
        return self.hmid - float(self.vol_dim[2] // 2)


def project(prob, vol, window=None, h=1.0, norm=None, proj=None):
    """vol[z,y,x] (local block) -> proj[t,h,w] += norm * h * sum_k T(src + k h u)."""
    vol, pv = _d(vol)
    nz, ny, nx = vol.shape
    nw, nh = (prob.proj_dim if window is None else window)
    if proj is None:
        proj = np.zeros((prob.n, int(nh), int(nw)))
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    norm, pn = _d(np.ones(prob.n) if norm is None else norm)
    org, po = _d(prob.origin)
    woff, pwo = _i(prob.woff)
    hoff, pho = _i(prob.hoff)
    g = [_d(a) for a in (prob.src, prob.det, prob.wstep, prob.hstep)]
    lib().orc_project(pv, nx, ny, nz, po, proj.ctypes.data_as(_dp), int(nw), int(nh), prob.n,
                      int(prob.proj_dim[0]), int(prob.proj_dim[1]), pwo, pho,
                      *[p for _, p in g], ctypes.c_double(h), pn)
    return proj


def adjoint(prob, proj, vol_shape, h=1.0, norm=None, vol=None):
    """Exact transpose of project(): vol += A^T proj."""
    proj, pp = _d(proj)
    _, nh, nw = proj.shape
    nz, ny, nx = vol_shape
    if vol is None:
        vol = np.zeros((nz, ny, nx))
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    norm, pn = _d(np.ones(prob.n) if norm is None else norm)
    org, po = _d(prob.origin)
    woff, pwo = _i(prob.woff)
    hoff, pho = _i(prob.hoff)
    g = [_d(a) for a in (prob.src, prob.det, prob.wstep, prob.hstep)]
    lib().orc_adjoint(vol.ctypes.data_as(_dp), nx, ny, nz, po, pp, nw, nh, prob.n,
                      int(prob.proj_dim[0]), int(prob.proj_dim[1]), pwo, pho,
                      *[p for _, p in g], ctypes.c_double(h), pn)
    return vol


def sample_count(prob, vol_shape, window=None, h=1.0):
    """Per-view number of samples inside the closed support box (algorithmic count)."""
    nz, ny, nx = vol_shape
    nw, nh = (prob.proj_dim if window is None else window)
    counts = np.zeros(prob.n, dtype=np.int64)
    org, po = _d(prob.origin)
    woff, pwo = _i(prob.woff)
    hoff, pho = _i(prob.hoff)
    g = [_d(a) for a in (prob.src, prob.det, prob.wstep, prob.hstep)]
    lib().orc_sample_count(nx, ny, nz, po, int(nw), int(nh), prob.n, int(prob.proj_dim[0]),
                           int(prob.proj_dim[1]), pwo, pho, *[p for _, p in g],
                           ctypes.c_double(h),
                           counts.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    return counts


def backproject(prob, proj, vol_shape, mode=2, norm=None, vol=None):
    """Voxel-driven backprojection: mode 2 FDK (U^2), 1 Katsevich (U^1), 0 unweighted general."""
    proj, pp = _d(proj)
    _, nh, nw = proj.shape
    nz, ny, nx = vol_shape
    if vol is None:
        vol = np.zeros((nz, ny, nx))
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    if norm is None:
        norm = np.full(prob.n, 1.0 / prob.n)
    norm, pn = _d(norm)
    org, po = _d(prob.origin)
    woff, pwo = _i(prob.woff)
    hoff, pho = _i(prob.hoff)
    if mode == 0:
        g = [_d(a) for a in (prob.src, prob.det, prob.wstep, prob.hstep)]
        lib().orc_backproject_general(vol.ctypes.data_as(_dp), nx, ny, nz, po, pp, nw, nh,
                                      prob.n, int(prob.proj_dim[0]), int(prob.proj_dim[1]),
                                      pwo, pho, *[p for _, p in g], pn)
    else:
        ang, pa = _d(prob.angles)
        zc, pz = _d(prob.zc)
        lib().orc_backproject_fdk(vol.ctypes.data_as(_dp), nx, ny, nz, po, pp, nw, nh, prob.n,
                                  int(prob.proj_dim[0]), int(prob.proj_dim[1]), pwo, pho, pa, pz,
                                  ctypes.c_double(prob.D), ctypes.c_double(prob.pw), pn,
                                  ctypes.c_int(1 if mode == 1 else 2))
    return vol


# ----------------------------------------------------------------------------------------------
# FDK pre-processing (fastcone.py:260-292), FP64 numpy restatements
# ----------------------------------------------------------------------------------------------
def ramp_filter_response(nw):
    """H[k], k = 0..Nw: pi * |fftfreq(2 Nw)| truncated to Nw+1 bins, H[0] = pi*0.125/(2Nw).

    fastcone.py:278-281 (np.resize keeps the first n/2+1 entries of |fftfreq(n)|).
    """
    n = 2 * nw
    f = np.abs(np.fft.fftfreq(n))
    f[0] = 0.125 * f[1]
    f *= np.pi
    return f[: n // 2 + 1].copy()


def rampfilter(proj):
    """fastcone.py:272-292: per view, zero-pad w to 2 Nw (data left-aligned), multiply the rfft by
    the ramp, irfft, crop [0, Nw). Returns a new float64 array."""
    proj = np.asarray(proj, dtype=np.float64)
    nt, nh, nw = proj.shape
    H = ramp_filter_response(nw)
    out = np.empty_like(proj)
    pad = np.zeros((nh, 2 * nw))
    for t in range(nt):
        pad[:] = 0.0
        pad[:, :nw] = proj[t]
        out[t] = np.fft.irfft(np.fft.rfft(pad, axis=1) * H, n=2 * nw, axis=1)[:, :nw]
    return out


def fbp_cosine_weight(shape, D, zorigin):
    """fastcone.py:261-267: D / sqrt(D^2 + hh^2 + ww^2), ww = j - Nw/2 (float division),
    hh = i - zOrigin[t], both in pixel units (no pixelWidth), per view."""
    nt, nh, nw = shape
    ww = np.arange(nw, dtype=np.float64) - nw / 2.0
    out = np.empty(shape)
    for t in range(nt):
        hh = np.arange(nh, dtype=np.float64)[:, None] - float(zorigin[t])
        out[t] = D / np.sqrt(D * D + hh * hh + ww[None, :] ** 2)
    return out


def fdk_cosine_weight(prob, window=None):
    """Physical-units FDK pre-weight of every window pixel: the cosine of the angle between the
    pixel's ray and the detector normal, |(det - src) . n| / |P - src| with n the unit normal of
    the detector plane spanned by wstep and hstep, P = det + gw wstep + gh hstep and
    gw = j + woff - Nw//2, gh = i + hoff - Nh//2 (the pixel convention of coneBeamGPU.cu:523-524).
    For the aligned geometry this is D / sqrt(D^2 + (gw pw)^2 + (gh pw)^2): the weight of
    fastcone.py:261-266 with the pixel offsets scaled by the pitch. Returns [n, nh, nw] FP64."""
    nw, nh = (prob.proj_dim if window is None else window)
    Nw, Nh = int(prob.proj_dim[0]), int(prob.proj_dim[1])
    out = np.empty((prob.n, int(nh), int(nw)))
    jj, ii = np.meshgrid(np.arange(int(nw)), np.arange(int(nh)))
    for t in range(prob.n):
        s, d, ws, hs = prob.src[t], prob.det[t], prob.wstep[t], prob.hstep[t]
        nrm = np.cross(ws, hs)
        nrm = nrm / np.linalg.norm(nrm)
        gw = jj + prob.woff[t] - Nw // 2
        gh = ii + prob.hoff[t] - Nh // 2
        P = d[None, None, :] + gw[..., None] * ws + gh[..., None] * hs
        out[t] = abs(float(np.dot(d - s, nrm))) / np.linalg.norm(P - s, axis=-1)
    return out


def fdk_filter(prob, proj, window=None, pw=None):
    """FDK pre-processing in physical units: cosine pre-weight (fdk_cosine_weight), then the
    reference's ramp (rampfilter, fastcone.py:272-292) divided by the pitch pw (default the
    problem's), so that FDK returns the attenuation per voxel. FP64."""
    pw = prob.pw if pw is None else float(pw)
    return rampfilter(np.asarray(proj, np.float64) * fdk_cosine_weight(prob, window)) / pw


# ----------------------------------------------------------------------------------------------
# Independent brute force for tiny problems (scipy order-1 splines, mode="grid-constant")
# ----------------------------------------------------------------------------------------------
def project_bruteforce(prob, vol, window=None, h=1.0, norm=None):
    from scipy.ndimage import map_coordinates
    vol = np.asarray(vol, dtype=np.float64)
    nz, ny, nx = vol.shape
    nw, nh = (prob.proj_dim if window is None else window)
    Nw, Nh = int(prob.proj_dim[0]), int(prob.proj_dim[1])
    norm = np.ones(prob.n) if norm is None else np.asarray(norm, np.float64)
    o = prob.origin
    out = np.zeros((prob.n, int(nh), int(nw)))
    # sample range: a generous fixed k range covering the whole bounding sphere of the block
    ext = np.array([nx, ny, nz], dtype=np.float64)
    centre = o + (ext - 1) / 2.0
    radius = 0.5 * np.linalg.norm(ext + 2.0) + 2.0
    jj, ii = np.meshgrid(np.arange(int(nw)), np.arange(int(nh)))
    for t in range(prob.n):
        s = prob.src[t]
        gw = jj + prob.woff[t] - Nw // 2
        gh = ii + prob.hoff[t] - Nh // 2
        P = prob.det[t][None, None, :] + gw[..., None] * prob.wstep[t] + gh[..., None] * prob.hstep[t]
        d = P - s
        u = d / np.linalg.norm(d, axis=-1, keepdims=True)
        dist = np.linalg.norm(centre - s)
        k0 = int(math.floor((dist - radius) / h)) - 1
        k1 = int(math.ceil((dist + radius) / h)) + 1
        ks = np.arange(k0, k1 + 1, dtype=np.float64) * h
        pts = (s - o)[None, None, None, :] + ks[:, None, None, None] * u[None]
        # map_coordinates wants [z, y, x] index order
        coords = np.stack([pts[..., 2], pts[..., 1], pts[..., 0]])
        vals = map_coordinates(vol, coords.reshape(3, -1), order=1, mode="grid-constant",
                               cval=0.0).reshape(pts.shape[:-1])
        out[t] = norm[t] * h * vals.sum(axis=0)
    return out


def backproject_bruteforce(prob, proj, vol_shape, mode=2, norm=None):
    """FDK/Katsevich/unweighted BP via a projective mapping built from the geometry vectors and
    scipy bilinear sampling; independent of the oracle's angle-based formulas."""
    from scipy.ndimage import map_coordinates
    proj = np.asarray(proj, dtype=np.float64)
    nt, nh, nw = proj.shape
    nz, ny, nx = vol_shape
    Nw, Nh = int(prob.proj_dim[0]), int(prob.proj_dim[1])
    norm = np.full(prob.n, 1.0 / prob.n) if norm is None else np.asarray(norm, np.float64)
    o = prob.origin
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([xx + o[0], yy + o[1], zz + o[2]], axis=-1).astype(np.float64)
    out = np.zeros(vol_shape)
    for t in range(nt):
        s, d, ws, hs = prob.src[t], prob.det[t], prob.wstep[t], prob.hstep[t]
        # solve s + lam (X - s) = d + w ws + h hs for (w, h, lam) per voxel
        r = X - s
        M = np.empty(r.shape[:-1] + (3, 3))
        M[..., :, 0] = ws
        M[..., :, 1] = hs
        M[..., :, 2] = -r
        rhs = np.broadcast_to(s - d, r.shape)
        sol = np.linalg.solve(M, rhs[..., None])[..., 0]
        w, hh, lam = sol[..., 0], sol[..., 1], sol[..., 2]
        wl = w + Nw // 2 - prob.woff[t]
        hl = hh + Nh // 2 - prob.hoff[t]
        vals = map_coordinates(proj[t], np.stack([hl.ravel(), wl.ravel()]), order=1,
                               mode="grid-constant", cval=0.0).reshape(w.shape)
        if mode == 0:
            wgt = 1.0
        else:
            # lam = |detector plane distance| / |voxel distance| along the ray = U for the
            # aligned geometry (the detector plane passes through the rotation axis)
            wgt = lam ** (2 if mode == 2 else 1)
        out += norm[t] * wgt * vals
    return out


# ----------------------------------------------------------------------------------------------
# Closed forms
# ----------------------------------------------------------------------------------------------
def ellipsoid_line_integrals(src, pts, ellipsoids):
    """Line integrals through a sum of uniform ellipsoids along rays from src through pts.

    src: [3]; pts: [..., 3]; ellipsoids: iterable of (centre[3], semi_axes[3], R[3,3], mu), with
    the ellipsoid {x : |diag(1/a) R^T (x - c)| <= 1}. The integral over the whole (infinite) line
    is mu times the chord length 2 sqrt(disc) / (u^T A u).
    """
    src = np.asarray(src, dtype=np.float64)
    d = np.asarray(pts, dtype=np.float64) - src
    u = d / np.linalg.norm(d, axis=-1, keepdims=True)
    out = np.zeros(u.shape[:-1])
    for c, a, R, mu in ellipsoids:
        A = R @ np.diag(1.0 / np.asarray(a, dtype=np.float64) ** 2) @ R.T
        q = src - np.asarray(c, dtype=np.float64)
        uA = u @ A
        qa = uA @ q
        aa = np.einsum("...i,...i->...", uA, u)
        cc = q @ A @ q - 1.0
        disc = qa * qa - aa * cc
        out += mu * np.where(disc > 0, 2.0 * np.sqrt(np.maximum(disc, 0.0)) / aa, 0.0)
    return out


# ----------------------------------------------------------------------------------------------
# Helical cone beam, cylindrical detector (Katsevich)
# ----------------------------------------------------------------------------------------------
class Helix:
    """a(l) = (R cos l, R sin l, zs(l)) in voxels; column j at fan angle (j - col_c) dg from
    e_v = -(cos l, sin l, 0) towards e_u = (-sin l, cos l, 0); row i at height (i - row_c) drow on
    the detector at distance SDD. hp = dzs/dl (voxels per radian), dl = view spacing."""

    def __init__(self, lambdas, zs, R, SDD, dg, drow, nw, nh, hp, col_c=None, row_c=None):
        self.lambdas = np.ascontiguousarray(lambdas, np.float64)
        self.zs = np.ascontiguousarray(np.broadcast_to(np.asarray(zs, np.float64),
                                                       self.lambdas.shape))
        self.R, self.SDD, self.dg, self.drow = float(R), float(SDD), float(dg), float(drow)
        self.nw, self.nh = int(nw), int(nh)
        self.hp = float(hp)
        self.col_c = (nw - 1) / 2.0 if col_c is None else float(col_c)
        self.row_c = (nh - 1) / 2.0 if row_c is None else float(row_c)
        self.dl = float(self.lambdas[1] - self.lambdas[0]) if self.lambdas.size > 1 else 1.0

    @property
    def n(self):
        return self.lambdas.size

    def ray_points(self, t):
        """Source a(l_t) [3] and one point on each pixel's ray [nh, nw, 3]."""
        l = self.lambdas[t]
        a = np.array([self.R * np.cos(l), self.R * np.sin(l), self.zs[t]])
        ev = np.array([-np.cos(l), -np.sin(l), 0.0])
        eu = np.array([-np.sin(l), np.cos(l), 0.0])
        g = (np.arange(self.nw) - self.col_c) * self.dg
        w = (np.arange(self.nh) - self.row_c) * self.drow
        d = (self.SDD * np.cos(g)[None, :, None] * ev + self.SDD * np.sin(g)[None, :, None] * eu
             + w[:, None, None] * np.array([0.0, 0.0, 1.0]))
        return a, a + d


def project_helical(hx, vol, origin, h=0.5):
    vol, pv = _d(vol)
    nz, ny, nx = vol.shape
    proj = np.zeros((hx.n, hx.nh, hx.nw))
    org, po = _d(origin)
    lam, pl = _d(hx.lambdas)
    zs, pz = _d(hx.zs)
    D = ctypes.c_double
    lib().orc_project_helical(pv, nx, ny, nz, po, proj.ctypes.data_as(_dp), hx.nw, hx.nh, hx.n,
                              pl, pz, D(hx.R), D(hx.SDD), D(hx.dg), D(hx.drow), D(hx.col_c),
                              D(hx.row_c), D(h))
    return proj


KATSEVICH_SIGN = 1.0


def katsevich(hx, proj, vol_shape, origin, npsi=None, sign=None, return_g5=False):
    """FP64 Katsevich reconstruction (see oracle.cpp: orc_katsevich)."""
    proj, pp = _d(proj)
    nz, ny, nx = vol_shape
    vol = np.zeros(vol_shape)
    npsi = 2 * hx.nh + 1 if npsi is None else int(npsi)
    org, po = _d(origin)
    lam, pl = _d(hx.lambdas)
    zs, pz = _d(hx.zs)
    g5 = np.zeros(proj.shape) if return_g5 else None
    D = ctypes.c_double
    lib().orc_katsevich(pp, hx.nw, hx.nh, hx.n, pl, pz, D(hx.R), D(hx.SDD), D(hx.dg),
                        D(hx.drow), D(hx.col_c), D(hx.row_c), D(hx.hp), D(hx.dl),
                        ctypes.c_int(npsi), D(KATSEVICH_SIGN if sign is None else sign),
                        vol.ctypes.data_as(_dp), nx, ny, nz, po,
                        g5.ctypes.data_as(_dp) if return_g5 else None)
    return (vol, g5) if return_g5 else vol
