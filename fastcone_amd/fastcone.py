"""Drop-in replacement for the reference's ``fastcone.py`` call surface.

Same function names, positional order, defaults and accumulate-into-caller-memory semantics as
the reference (fastcone.py:52-292), backed by libFastcone.so through the same five ``GPU_CB_*``
C entry points. Differences, all deliberate:

* bad arguments raise ``ValueError`` / ``TypeError`` instead of ``sys.exit`` (fastcone.py:75-82),
  and a failed library call raises ``FastconeError`` (the reference checks nothing);
* ``None`` defaults are tested with ``is None`` (the reference's ``== None`` fails on arrays with
  numpy 2);
* the arrays handed to C must be C-contiguous float32 (the reference silently passes strided
  buffers through ``ctypes.data_as``);
* ``rampfilter`` filters every view in one vectorised FFT instead of a per-view loop, with the
  same FP64 arithmetic and the same filter.

Usage: ``from fastcone_amd import fastcone as cone``.
"""
import ctypes
import math

import numpy as np

from . import _native

# Data types compatible with c_float and c_int (fastcone.py:49-50)
FLOAT_TYPE = np.float32
INT_TYPE = np.int32

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def _need_f32_c(name, a):
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name} must be a numpy array")
    if a.dtype != FLOAT_TYPE:
        raise TypeError(f"{name} array does not have data type np.float32")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name} array must be C-contiguous")


def _common_checks(fn, vol, proj, angle, zOrigin):
    _need_f32_c("Volume", vol)
    _need_f32_c("Projection", proj)
    if vol.ndim != 3 or proj.ndim != 3:
        raise ValueError(f"{fn}: vol and proj must be 3-D ([z,y,x] and [theta,h,w])")
    angle = np.asarray(angle)
    zOrigin = np.asarray(zOrigin)
    if angle.size != proj.shape[0]:
        raise ValueError("Number of angles supplied does not match number of projections in "
                         "proj array")
    if zOrigin.size != proj.shape[0]:
        raise ValueError("Number of zOrigin elements does not match number of projections in "
                         "proj array")
    return angle, zOrigin


def _as(a, dtype):
    return np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=dtype)


def _call(fname, GPUID, vol, proj, angle, hmid, sourceDistance, pixelWidth, volDims, volStart,
          volEnd, projDims, projWHSize, projWOffset, projHOffset, norm, last):
    L = _native.lib()
    L.fc_clear_error()
    keep = [angle, hmid, volDims, volStart, volEnd, projDims, projWHSize, projWOffset,
            projHOffset, norm]
    args = [ctypes.c_int(int(GPUID)), vol.ctypes.data_as(_fp), proj.ctypes.data_as(_fp),
            angle.ctypes.data_as(_fp), hmid.ctypes.data_as(_fp),
            ctypes.c_float(float(sourceDistance)), ctypes.c_float(float(pixelWidth)),
            volDims.ctypes.data_as(_ip), volStart.ctypes.data_as(_ip),
            volEnd.ctypes.data_as(_ip), projDims.ctypes.data_as(_ip),
            projWHSize.ctypes.data_as(_ip), projWOffset.ctypes.data_as(_ip),
            projHOffset.ctypes.data_as(_ip), norm.ctypes.data_as(_fp)]
    if isinstance(last, np.ndarray):
        keep.append(last)
        args.append(last.ctypes.data_as(_fp))
    else:
        args.append(ctypes.c_int(int(last)))
    getattr(L, fname)(*args)
    del keep
    _native.check()


def project(vol, proj, angle, sourceDistance, zOrigin, volStart=np.array([0, 0, 0], dtype=INT_TYPE),
            volDims=None, projWOffset=None, projHOffset=None, projDims=None, norm=None, GPUID=0,
            misAlignments=np.array([0, 0, 0, 0, 0, 0], dtype=FLOAT_TYPE), pixelWidth=1.0):
    """Forward-project vol[z,y,x] into proj[theta,h,w], adding to what proj holds.

    Geometry as the reference (fastcone.py:52-73): cone beam, angles in radians, sourceDistance
    from the source to the rotation axis in voxels, zOrigin[t] the z index (from volume element
    0) of the source and detector centre, pixelWidth the demagnified detector pitch in voxels,
    misAlignments [Dw, Dh, Dl (pixels), Dphi, Dtheta, Dpsi (radians)]. Rays are sampled at unit
    steps from the source. Defaults: norm = 1, volDims = vol.shape, offsets 0. After the call the
    reference's centring quirk is reproduced: where projWOffset[t] <= 0, columns up to global
    column 0 are zeroed (fastcone.py:140-143). Returns 1.
    """
    angle, zOrigin = _common_checks("project", vol, proj, angle, zOrigin)
    volStart = np.asarray(volStart)
    volEnd = np.array([vol.shape[0] + volStart[0], vol.shape[1] + volStart[1],
                       vol.shape[2] + volStart[2]], dtype=INT_TYPE)
    if volDims is None:
        volDims = np.array([vol.shape[0], vol.shape[1], vol.shape[2]], dtype=INT_TYPE)
    if projWOffset is None:
        projWOffset = np.zeros(angle.size, dtype=INT_TYPE)
    if projHOffset is None:
        projHOffset = np.zeros(angle.size, dtype=INT_TYPE)
    projWOffset = _as(projWOffset, INT_TYPE)
    projHOffset = _as(projHOffset, INT_TYPE)
    if projWOffset.size != angle.size or projHOffset.size != angle.size:
        raise ValueError("projWOffset / projHOffset must have one entry per projection")
    if projDims is None:
        projDims = np.array([proj.shape[0], proj.shape[1] + projHOffset.max(initial=0),
                             proj.shape[2] + projWOffset.max(initial=0)], dtype=INT_TYPE)
    if norm is None:
        norm = np.ones(angle.size, dtype=FLOAT_TYPE)
    norm = _as(norm, FLOAT_TYPE)
    if norm.size != angle.size:
        raise ValueError("norm must have one entry per projection")
    projDims = np.array(projDims, dtype=INT_TYPE).copy()
    projDims[0] = proj.shape[0]  # projDims[0] means something different to the C code
    projWHSize = np.array([proj.shape[1], proj.shape[2]], dtype=INT_TYPE)
    hMidplane = _as(zOrigin, FLOAT_TYPE)
    misAlignments = _as(misAlignments, FLOAT_TYPE)
    if misAlignments.size != 6:
        raise ValueError("misAlignments must have 6 elements")
    # reverse to the C order [x,y,z] / [w,h,t]
    volStart = _as(np.asarray(volStart)[::-1], INT_TYPE)
    volEnd = _as(volEnd[::-1], INT_TYPE)
    volDims = _as(np.asarray(volDims)[::-1], INT_TYPE)
    projDims = _as(projDims[::-1], INT_TYPE)
    projWHSize = _as(projWHSize[::-1], INT_TYPE)
    _call("GPU_CB_misalign_project", GPUID, vol, proj, _as(angle, FLOAT_TYPE), hMidplane,
          sourceDistance, pixelWidth, volDims, volStart, volEnd, projDims, projWHSize,
          projWOffset, projHOffset, norm, misAlignments)
    # The reference centres the detector at pixel N/2 instead of N/2 - 1/2 and zeroes column 0
    # to compensate (fastcone.py:140-143).
    for tt in range(proj.shape[0]):
        if projWOffset[tt] <= 0:
            proj[tt, :, 0:(1 + projWOffset[tt])] = 0.0
    return 1


def backproject(vol, proj, angle, sourceDistance, zOrigin, volStart=np.array([0, 0, 0], dtype=INT_TYPE),
                volDims=None, projWOffset=None, projHOffset=None, projDims=None, norm=None,
                GPUID=0, kernelMode=2, misAlignments=None, pixelWidth=1.0):
    """Backproject proj[theta,h,w] into vol[z,y,x], adding to what vol holds.

    kernelMode 2 = Feldkamp weight U^2, 1 = Katsevich weight U, 0 = unweighted general geometry
    (the only mode that accepts misAlignments), fastcone.py:147-256. Defaults:
    norm = 1/projDims[0], volDims = volStart + vol.shape. Returns 1.
    """
    angle, zOrigin = _common_checks("backproject", vol, proj, angle, zOrigin)
    if misAlignments is not None and kernelMode != 0:
        raise ValueError("Cannot conduct misaligned backprojection unless kernel mode is set to "
                         "0 (i.e. unweighted, not Feldkamp or Katsevich).")
    volStart = np.asarray(volStart)
    volEnd = np.array([vol.shape[0] + volStart[0], vol.shape[1] + volStart[1],
                       vol.shape[2] + volStart[2]], dtype=INT_TYPE)
    if volDims is None:
        volDims = volEnd.copy()
    if projWOffset is None:
        projWOffset = np.zeros(angle.size, dtype=INT_TYPE)
    if projHOffset is None:
        projHOffset = np.zeros(angle.size, dtype=INT_TYPE)
    projWOffset = _as(projWOffset, INT_TYPE)
    projHOffset = _as(projHOffset, INT_TYPE)
    if projWOffset.size != angle.size or projHOffset.size != angle.size:
        raise ValueError("projWOffset / projHOffset must have one entry per projection")
    if projDims is None:
        projDims = np.array([proj.shape[0], proj.shape[1] + projHOffset.max(initial=0),
                             proj.shape[2] + projWOffset.max(initial=0)], dtype=INT_TYPE)
    if norm is None:
        norm = np.ones(angle.size, dtype=FLOAT_TYPE) / projDims[0]
    norm = _as(norm, FLOAT_TYPE)
    if norm.size != angle.size:
        raise ValueError("norm must have one entry per projection")
    projDims = np.array(projDims, dtype=INT_TYPE).copy()
    projDims[0] = proj.shape[0]
    projWHSize = np.array([proj.shape[1], proj.shape[2]], dtype=INT_TYPE)
    hMidplane = _as(zOrigin, FLOAT_TYPE)
    volStart = _as(np.asarray(volStart)[::-1], INT_TYPE)
    volEnd = _as(volEnd[::-1], INT_TYPE)
    volDims = _as(np.asarray(volDims)[::-1], INT_TYPE)
    projDims = _as(projDims[::-1], INT_TYPE)
    projWHSize = _as(projWHSize[::-1], INT_TYPE)
    if misAlignments is None:
        _call("GPU_CB_K_backproject", GPUID, vol, proj, _as(angle, FLOAT_TYPE), hMidplane,
              sourceDistance, pixelWidth, volDims, volStart, volEnd, projDims, projWHSize,
              projWOffset, projHOffset, norm, int(kernelMode))
    else:
        misAlignments = _as(misAlignments, FLOAT_TYPE)
        if misAlignments.size != 6:
            raise ValueError("misAlignments must have 6 elements")
        _call("GPU_CB_misalign_backproject", GPUID, vol, proj, _as(angle, FLOAT_TYPE),
              hMidplane, sourceDistance, pixelWidth, volDims, volStart, volEnd, projDims,
              projWHSize, projWOffset, projHOffset, norm, misAlignments)
    return 1


def fbp(vol, proj, projAngles, specimenLength, zOrigin, volStart=np.array([0, 0, 0], dtype=INT_TYPE),
        volDims=None, GPUID=0, pixelWidth=1.0):
    """FDK: cosine pre-weighting, ramp filter, Feldkamp backprojection (fastcone.py:260-270).

    The cosine weight reproduces the reference exactly, including its use of pixel-index offsets
    without pixelWidth and of the volume index zOrigin as the detector row of the source
    (ww = j - Nw/2, hh = i - zOrigin[t]). proj is modified in place. Returns the last view's
    weight array, as the reference does.
    """
    ww = np.arange(0.0, proj.shape[2], 1.0) - proj.shape[2] / 2.0
    ww = ww.reshape(1, proj.shape[2])
    zOrigin = np.asarray(zOrigin)
    reScale = None
    for pp in range(proj.shape[0]):
        hh = (np.arange(0.0, proj.shape[1], 1.0) - float(zOrigin[pp])).reshape(proj.shape[1], 1)
        reScale = specimenLength / np.sqrt(specimenLength * specimenLength + hh * hh + ww * ww)
        proj[pp, :, :] = proj[pp, :, :] * reScale[:, :]
    rampfilter(proj)
    backproject(vol, proj, projAngles, specimenLength, zOrigin, volStart, volDims, GPUID=GPUID,
                pixelWidth=pixelWidth)
    return reScale


def ramp_response(nw, window=None, scale=0.125):
    """Ramp of fastcone.py:278-281 on the zero-padded length n = 2 Nw: H[k] = pi k / n for
    k = 1..Nw and H[0] = pi * 0.125 / n (rfft bins). window "cos" multiplies by cos(pi f) and
    "gauss" by exp(-f^2 / (2 scale^2)), f = k / n in cycles per pixel (the windows the reference's
    cosrampfilter / gaussrampfilter aim at, fastcone.py:294-319)."""
    n = 2 * int(nw)
    f = np.arange(n // 2 + 1, dtype=np.float64) / n
    H = np.pi * f
    H[0] = np.pi * 0.125 / n
    if window == "cos":
        H *= np.cos(np.pi * f)
    elif window == "gauss":
        H *= np.exp(-0.5 * (f / float(scale)) ** 2)
    elif window is not None:
        raise ValueError(f"unknown ramp window {window!r}")
    return H


def _filter_views(proj, H):
    nt, nh, nw = proj.shape
    chunk = max(1, (1 << 24) // max(1, nh * nw))  # bound the FP64 temporaries
    for t0 in range(0, nt, chunk):
        t1 = min(nt, t0 + chunk)
        spec = np.fft.rfft(proj[t0:t1].astype(np.float64), n=2 * nw, axis=2)
        spec *= H
        proj[t0:t1] = np.fft.irfft(spec, n=2 * nw, axis=2)[..., :nw]


def rampfilter(proj):
    """Ramp-filter every view along w in place (fastcone.py:272-292): zero-pad to 2 Nw with the
    data left-aligned, multiply the FP64 rfft by the ramp, inverse transform, crop [0, Nw)."""
    _filter_views(proj, ramp_response(proj.shape[2]))
    return


def cosrampfilter(proj):
    """Cosine-windowed ramp filter, in place (fastcone.py:294-305, fixed). The reference does not
    run as written: it transforms back only after its loop over views (so one view is filtered)
    and scales its cosine by 1/Nw (so the window is ~1). Here every view gets rampfilter's padded
    ramp times cos(pi f), which falls to zero at the Nyquist frequency."""
    _filter_views(proj, ramp_response(proj.shape[2], "cos"))
    return


def gaussrampfilter(proj, scale=0.125):
    """Gaussian-windowed ramp filter, in place (fastcone.py:307-319, fixed: the reference uses
    scipy without importing it and transforms back only after its loop). The window is the
    reference's Gaussian over frequency indices with standard deviation scale * Nw, i.e.
    exp(-f^2 / (2 scale^2)) in cycles per pixel, on rampfilter's padded ramp."""
    _filter_views(proj, ramp_response(proj.shape[2], "gauss", scale))
    return


def set_tile_limits(max_voxels=0, max_views=0):
    """Out-of-core tiling for the host-memory calls: z slabs of at most max_voxels voxels and view
    batches of at most max_views views per device pass (0 = whole)."""
    L = _native.lib()
    L.fc_clear_error()
    L.fc_set_tile_limits(int(max_voxels), int(max_views))
    _native.check()
