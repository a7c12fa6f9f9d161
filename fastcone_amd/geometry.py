"""Cone-beam geometry descriptions for the device-pointer API.

A ``Geometry`` holds the FP64 per-view records built by libFastcone (fc_views_flat /
fc_views_cylinder), the volume block it applies to and the detector window shape. All lengths
are in voxels; a voxel with global index g along an axis of N voxels has coordinate g - N//2, as
in the reference (coneBeamGPU.cu:597-599). Python shapes are C order: volume [nz, ny, nx],
projections [nviews, nh, nw].
"""
import numpy as np

from . import _native


class Geometry:
    def __init__(self, views, vol_shape, origin, det_shape, kind="flat"):
        self.views = np.ascontiguousarray(views, dtype=np.float64).reshape(-1, _native.FC_VIEW_DOUBLES)
        self.vol_shape = tuple(int(v) for v in vol_shape)      # (nz, ny, nx)
        self.origin = np.asarray(origin, dtype=np.float64)     # coordinate of local voxel 0 (x,y,z)
        self.det_shape = tuple(int(v) for v in det_shape)      # (nh, nw)
        self.kind = kind
        self._dev = {}

    @property
    def nviews(self):
        return self.views.shape[0]

    @property
    def proj_shape(self):
        return (self.nviews,) + self.det_shape

    def views_on(self, device):
        """Device copy of the records (cached per device)."""
        import torch
        key = str(device)
        t = self._dev.get(key)
        if t is None:
            t = torch.from_numpy(self.views).to(device)
            self._dev[key] = t
        return t

    def subset(self, idx):
        """Geometry restricted to the views `idx` (slice or index array)."""
        return Geometry(self.views[idx], self.vol_shape, self.origin, self.det_shape, self.kind)

    def with_volume(self, vol_shape, origin):
        return Geometry(self.views, vol_shape, origin, self.det_shape, self.kind)

    # ---- constructors ----------------------------------------------------------------------
    @classmethod
    def circular(cls, angles, D, pw, vol_shape, det_shape, hmid=None, vol_dim=None,
                 vol_start=(0, 0, 0), woff=None, hoff=None, misalign=None, det_dim=None,
                 pw_h=None):
        """Reference geometry (coneBeamGPU.cu:1672-1734): source at distance D from the axis,
        flat virtual detector through the axis with pitch pw (voxels), detector centre at global
        pixel N//2. vol_shape/det_shape are the local block / window in C order; vol_dim and
        det_dim (C order) default to them; vol_start (z,y,x) and woff/hoff place the block and the
        windows; hmid is the source z index from volume element 0 (default vol_dim_z / 2)."""
        angles = np.ascontiguousarray(angles, dtype=np.float64).reshape(-1)
        n = angles.size
        vol_shape = tuple(int(v) for v in vol_shape)
        vol_dim = tuple(int(v) for v in (vol_dim if vol_dim is not None else vol_shape))
        det_shape = tuple(int(v) for v in det_shape)
        det_dim = tuple(int(v) for v in (det_dim if det_dim is not None else det_shape))
        if hmid is None:
            hmid = np.full(n, vol_dim[0] / 2.0)
        hmid = np.ascontiguousarray(np.broadcast_to(np.asarray(hmid, np.float64), (n,)))
        woff = np.zeros(n, np.int32) if woff is None else np.ascontiguousarray(woff, np.int32)
        hoff = np.zeros(n, np.int32) if hoff is None else np.ascontiguousarray(hoff, np.int32)
        mis = None if misalign is None else np.ascontiguousarray(misalign, np.float64)
        views = np.zeros((n, _native.FC_VIEW_DOUBLES), np.float64)
        L = _native.lib()
        pd = _native.int_array([det_dim[1], det_dim[0]])
        rc = L.fc_views_flat(
            n, angles.ctypes.data_as(_native._c_dbl_p), hmid.ctypes.data_as(_native._c_dbl_p),
            float(D), float(pw), float(pw if pw_h is None else pw_h),
            None if mis is None else mis.ctypes.data_as(_native._c_dbl_p), int(vol_dim[0]), pd,
            woff.ctypes.data_as(_native._c_int_p), hoff.ctypes.data_as(_native._c_int_p),
            views.ctypes.data_as(_native._c_dbl_p))
        _native.check(rc)
        vs = np.asarray(vol_start, dtype=np.int64)  # (z, y, x)
        vd = np.asarray(vol_dim, dtype=np.int64)
        origin = np.array([vs[2] - vd[2] // 2, vs[1] - vd[1] // 2, vs[0] - vd[0] // 2],
                          dtype=np.float64)
        return cls(views, vol_shape, origin, det_shape, "flat")

    @classmethod
    def cylinder(cls, lambdas, src_z, R, SDD, dgamma, drow, vol_shape, det_shape,
                 col_c=None, row_c=None, vol_dim=None, vol_start=(0, 0, 0)):
        """Curved (cylindrical, source-centred) detector on a circular or helical source path
        a(l) = (R cos l, R sin l, src_z) in voxels. Column j has fan angle (j - col_c) * dgamma
        (radians), row i height (i - row_c) * drow at the detector. col_c / row_c default to the
        symmetric centres (nw - 1) / 2 and (nh - 1) / 2."""
        lambdas = np.ascontiguousarray(lambdas, dtype=np.float64).reshape(-1)
        n = lambdas.size
        vol_shape = tuple(int(v) for v in vol_shape)
        vol_dim = tuple(int(v) for v in (vol_dim if vol_dim is not None else vol_shape))
        det_shape = tuple(int(v) for v in det_shape)
        nh, nw = det_shape
        col_c = np.full(n, (nw - 1) / 2.0) if col_c is None else np.broadcast_to(
            np.asarray(col_c, np.float64), (n,))
        row_c = np.full(n, (nh - 1) / 2.0) if row_c is None else np.broadcast_to(
            np.asarray(row_c, np.float64), (n,))
        col_c = np.ascontiguousarray(col_c, dtype=np.float64)
        row_c = np.ascontiguousarray(row_c, dtype=np.float64)
        src_z = np.ascontiguousarray(np.broadcast_to(np.asarray(src_z, np.float64), (n,)))
        views = np.zeros((n, _native.FC_VIEW_DOUBLES), np.float64)
        L = _native.lib()
        rc = L.fc_views_cylinder(
            n, lambdas.ctypes.data_as(_native._c_dbl_p), src_z.ctypes.data_as(_native._c_dbl_p),
            float(R), float(SDD), float(dgamma), float(drow),
            col_c.ctypes.data_as(_native._c_dbl_p), row_c.ctypes.data_as(_native._c_dbl_p),
            views.ctypes.data_as(_native._c_dbl_p))
        _native.check(rc)
        vs = np.asarray(vol_start, dtype=np.int64)
        vd = np.asarray(vol_dim, dtype=np.int64)
        origin = np.array([vs[2] - vd[2] // 2, vs[1] - vd[1] // 2, vs[0] - vd[0] // 2],
                          dtype=np.float64)
        g = cls(views, vol_shape, origin, det_shape, "cylinder")
        g.helix = dict(lambdas=lambdas, src_z=src_z, R=float(R), SDD=float(SDD),
                       dgamma=float(dgamma), drow=float(drow), col_c=col_c, row_c=row_c)
        return g
