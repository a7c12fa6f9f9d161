"""Single-GPU driver replacing the reference's mpi_fastcone.py.

The reference splits a 4-D volume [t, z, y, x] into chunks (default [1, 512, 512, 512],
mpi_fastcone.py:194), finds for each chunk the views that see it and each view's detector window by
projecting the chunk's corners (mpi_fastcone.py:234-368), ships chunks and windows between MPI
ranks, and calls fastcone.project / backproject per chunk (mpi_fastcone.py:371-526). Here one GPU
processes its whole share at its real shape, with no MPI fan-out and no collectives: the same chunk
and window planning, then one projector / backprojector call per chunk through the drop-in
fastcone surface, accumulating into the caller's arrays.

Differences from the reference planning, all deliberate:
  * windows are computed from libFastcone's FP64 projective matrices for the chunk's interpolation
    support (the chunk grown by one voxel) plus one pixel, so chunked results equal unchunked ones
    (the reference projects the bare corners in FP32 numpy without the pixel width,
    mpi_fastcone.py:315-326, and checks the h range against the detector width,
    mpi_fastcone.py:276-277);
  * `comm` is accepted for signature compatibility and must be None or of size 1.
"""
import numpy as np

from . import fastcone as cone
from .geometry import Geometry

FLOAT_TYPE = np.float32
INT_TYPE = np.int32


class ChunkPlan:
    """Chunks of a local 4-D volume block [t, z, y, x] (mpi_fastcone.infoVol, one rank)."""

    def __init__(self, local_dims, local_start, total_dims, chunk=(1, 512, 512, 512)):
        self.local_dims = np.asarray(local_dims, dtype=np.int64)
        self.local_start = np.asarray(local_start, dtype=np.int64)
        self.dims = np.asarray(total_dims, dtype=np.int64)
        self.chunk_dims = np.minimum(np.asarray(chunk, dtype=np.int64), self.local_dims)
        n = -(-self.local_dims // self.chunk_dims)  # ceil
        grid = np.indices(tuple(n)).reshape(4, -1).T
        self.local_offset = grid * self.chunk_dims
        self.offset = self.local_offset + self.local_start
        # chunk extents (the last chunk along an axis may be short)
        self.extent = np.minimum(self.chunk_dims, self.local_dims - self.local_offset)

    def __len__(self):
        return self.offset.shape[0]


class WindowPlan:
    """Per chunk: the views that see it and their detector windows (mpi_fastcone.infoGeometry)."""

    def __init__(self, plan, angles, zsource, sample_distance, proj_dims, pixel_width=1.0,
                 misalign=None, proj_stride=1, proj_offset=0):
        angles = np.asarray(angles, dtype=np.float64)
        zsource = np.asarray(zsource, dtype=np.float64)
        nv = angles.size
        Nh, Nw = int(proj_dims[1]), int(proj_dims[2])
        dims = plan.dims[1:4]  # z, y, x
        self.views, self.woff, self.hoff, self.wsize, self.hsize = [], [], [], [], []
        # FP64 records of the whole detector for every view
        g = Geometry.circular(angles, sample_distance, pixel_width, tuple(dims), (Nh, Nw),
                              hmid=zsource, vol_dim=tuple(dims), misalign=misalign)
        M = g.views[:, 12:24].reshape(nv, 3, 4)
        for c in range(len(plan)):
            lo = plan.offset[c, 1:4] - dims // 2 - 1.0          # z, y, x coordinates
            hi = lo + plan.extent[c, 1:4] + 1.0
            corners = np.array([[x, y, z, 1.0] for z in (lo[0], hi[0]) for y in (lo[1], hi[1])
                                for x in (lo[2], hi[2])])
            abc = np.einsum("tij,kj->tki", M, corners)
            u = abc[..., 0] / abc[..., 2]
            v = abc[..., 1] / abc[..., 2]
            w0 = np.floor(u.min(1)) - 1
            w1 = np.ceil(u.max(1)) + 1
            h0 = np.floor(v.min(1)) - 1
            h1 = np.ceil(v.max(1)) + 1
            ok = (w1 >= 0) & (w0 <= Nw - 1) & (h1 >= 0) & (h0 <= Nh - 1)
            ok &= ((np.arange(nv) + proj_offset) % proj_stride) == 0
            idx = np.nonzero(ok)[0]
            w0c = np.clip(w0[idx], 0, Nw - 1).astype(INT_TYPE)
            w1c = np.clip(w1[idx], 0, Nw - 1).astype(INT_TYPE)
            h0c = np.clip(h0[idx], 0, Nh - 1).astype(INT_TYPE)
            h1c = np.clip(h1[idx], 0, Nh - 1).astype(INT_TYPE)
            self.views.append(idx)
            self.woff.append(w0c)
            self.hoff.append(h0c)
            self.wsize.append(w1c - w0c + 1)
            self.hsize.append(h1c - h0c + 1)


def _check_comm(comm):
    if comm is not None and getattr(comm, "size", 1) != 1:
        raise ValueError("the single-GPU driver runs one GPU's share: comm must be None or of "
                         "size 1 (multi-GPU sharding is not part of this build)")


def project(plan, wplan, vol, proj, angles, zsource, sample_distance, pixel_width=1.0,
            misalign=None, GPUID=0):
    """proj[theta, h, w] += forward projection of every chunk of vol[t, z, y, x] (t = 0 block)."""
    mis = np.zeros(6, FLOAT_TYPE) if misalign is None else np.asarray(misalign, FLOAT_TYPE)
    angles = np.asarray(angles, FLOAT_TYPE)
    zsource = np.asarray(zsource, FLOAT_TYPE)
    for c in range(len(plan)):
        idx = wplan.views[c]
        if idx.size == 0:
            continue
        lo, ex = plan.local_offset[c], plan.extent[c]
        for t in range(ex[0]):
            vol_c = np.ascontiguousarray(vol[lo[0] + t, lo[1]:lo[1] + ex[1], lo[2]:lo[2] + ex[2],
                                             lo[3]:lo[3] + ex[3]], dtype=FLOAT_TYPE)
            hs, ws = int(wplan.hsize[c].max()), int(wplan.wsize[c].max())
            pw = np.zeros((idx.size, hs, ws), FLOAT_TYPE)
            cone.project(vol_c, pw, angles[idx], sample_distance, zsource[idx],
                         volStart=plan.offset[c, 1:4].astype(INT_TYPE),
                         volDims=plan.dims[1:4].astype(INT_TYPE), projWOffset=wplan.woff[c],
                         projHOffset=wplan.hoff[c],
                         projDims=np.array([idx.size, proj.shape[1], proj.shape[2]], INT_TYPE),
                         misAlignments=mis, GPUID=GPUID, pixelWidth=pixel_width)
            # add each view's window back (the reference's receive-and-add, mpi_fastcone.py:437-447)
            for q, v in enumerate(idx):
                h0, w0 = wplan.hoff[c][q], wplan.woff[c][q]
                hh, wwid = wplan.hsize[c][q], wplan.wsize[c][q]
                proj[v, h0:h0 + hh, w0:w0 + wwid] += pw[q, :hh, :wwid]
    return 1


def backproject(plan, wplan, vol, proj, angles, zsource, sample_distance, pixel_width=1.0,
                misalign=None, kernel_mode=2, GPUID=0, total_views=None):
    """vol[t, z, y, x] += backprojection of proj into every chunk (norm 1/total views)."""
    angles = np.asarray(angles, FLOAT_TYPE)
    zsource = np.asarray(zsource, FLOAT_TYPE)
    nv_total = proj.shape[0] if total_views is None else int(total_views)
    for c in range(len(plan)):
        idx = wplan.views[c]
        if idx.size == 0:
            continue
        lo, ex = plan.local_offset[c], plan.extent[c]
        hs, ws = int(wplan.hsize[c].max()), int(wplan.wsize[c].max())
        pw = np.zeros((idx.size, hs, ws), FLOAT_TYPE)
        for q, v in enumerate(idx):
            h0, w0 = wplan.hoff[c][q], wplan.woff[c][q]
            hh, wwid = wplan.hsize[c][q], wplan.wsize[c][q]
            pw[q, :hh, :wwid] = proj[v, h0:h0 + hh, w0:w0 + wwid]
        norm = np.full(idx.size, 1.0 / nv_total, FLOAT_TYPE)
        for t in range(ex[0]):
            vol_c = np.zeros((ex[1], ex[2], ex[3]), FLOAT_TYPE)
            cone.backproject(vol_c, pw, angles[idx], sample_distance, zsource[idx],
                             volStart=plan.offset[c, 1:4].astype(INT_TYPE),
                             volDims=plan.dims[1:4].astype(INT_TYPE), projWOffset=wplan.woff[c],
                             projHOffset=wplan.hoff[c],
                             projDims=np.array([idx.size, proj.shape[1], proj.shape[2]], INT_TYPE),
                             norm=norm, GPUID=GPUID, kernelMode=kernel_mode,
                             misAlignments=None if misalign is None else
                             np.asarray(misalign, FLOAT_TYPE),
                             pixelWidth=pixel_width)
            vol[lo[0] + t, lo[1]:lo[1] + ex[1], lo[2]:lo[2] + ex[2], lo[3]:lo[3] + ex[3]] += vol_c
    return 1


def simpleProject(comm, proj_P, vol_V, angles, volOffset, volDims, zSourceIndex, sampleDistance,
                  chunk=np.array([1, 512, 512, 512], dtype=INT_TYPE),
                  misAlignments=np.array([0, 0, 0, 0, 0, 0], dtype=FLOAT_TYPE), pixelWidth=1.0,
                  GPUID=0, projStride=1, projOffset=0):
    """mpi_fastcone.simpleProject on one GPU (mpi_fastcone.py:528-553): adds the projection of
    vol_V[t, z, y, x] (global offset volOffset within volDims) to proj_P[theta, h, w]."""
    _check_comm(comm)
    plan = ChunkPlan(vol_V.shape, volOffset, volDims, chunk)
    wplan = WindowPlan(plan, angles, zSourceIndex, sampleDistance, proj_P.shape, pixelWidth,
                       misAlignments, projStride, projOffset)
    return project(plan, wplan, vol_V, proj_P, angles, zSourceIndex, sampleDistance, pixelWidth,
                   misAlignments, GPUID)


def simpleBackproject(comm, proj_P, vol_V, angles, volOffset, volDims, zSourceIndex,
                      sampleDistance, chunk=np.array([1, 512, 512, 512], dtype=INT_TYPE),
                      misAlignments=None, kernelMode=2, pixelWidth=1.0, GPUID=0, projStride=1,
                      projOffset=0):
    """mpi_fastcone.simpleBackproject on one GPU (mpi_fastcone.py:556-591): adds the
    backprojection of proj_P (norm 1/number of views) to vol_V[t, z, y, x]."""
    _check_comm(comm)
    if misAlignments is not None and kernelMode != 0:
        raise ValueError("misaligned backprojection needs kernelMode 0")
    plan = ChunkPlan(vol_V.shape, volOffset, volDims, chunk)
    wplan = WindowPlan(plan, angles, zSourceIndex, sampleDistance, proj_P.shape, pixelWidth,
                       misAlignments, projStride, projOffset)
    return backproject(plan, wplan, vol_V, proj_P, angles, zSourceIndex, sampleDistance,
                       pixelWidth, misAlignments, kernelMode, GPUID, total_views=proj_P.shape[0])


def createVol(volDims, volLocalDims=None):
    """mpi_fastcone.createMpiVol on one GPU (mpi_fastcone.py:593-639): the whole volume lives on
    this GPU; returns (vol_V, volOffset)."""
    volDims = np.asarray(volDims, dtype=INT_TYPE)
    if volLocalDims is not None and np.any(np.asarray(volLocalDims) != volDims):
        raise ValueError("one GPU holds the whole volume: volLocalDims must equal volDims")
    return np.zeros(volDims, dtype=FLOAT_TYPE), np.zeros(4, dtype=INT_TYPE)
