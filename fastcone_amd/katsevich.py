"""Helical cone-beam CT on a cylindrical detector: geometry and exact Katsevich reconstruction.

The reference's "Katsevich" mode (kernelMode 1, coneBeamGPU.cu:258-338) is only a U-weighted
backprojection of data the caller must pre-filter, on a flat detector with one pitch. Here the
whole algorithm runs on the device in libFastcone (fc_katsevich_dev): fixed-direction
derivative, cone-angle weight, kappa-line forward rebinning, 1/sin Hilbert filtering, backward
rebinning and backprojection over each voxel's PI interval (Tam-Danielsson window).
"""
import math

import numpy as np
import torch

from . import _native
from .geometry import Geometry
from .ops import _dtype_code, _need_device, _stream


class HelixScan:
    """Helix a(l) = (R cos l, R sin l, zs0 + hp (l - l0)), l_k = l0 + k dl (voxel units), and a
    cylindrical detector of nh x nw pixels: column pitch dgamma (radians), row pitch drow (at the
    detector), source-to-detector SDD, symmetric centres."""

    def __init__(self, nviews, dl, l0, zs0, hp, R, SDD, dgamma, drow, nw, nh, col_c=None,
                 row_c=None):
        self.nviews, self.dl, self.l0, self.zs0, self.hp = int(nviews), float(dl), float(l0), \
            float(zs0), float(hp)
        self.R, self.SDD, self.dgamma, self.drow = float(R), float(SDD), float(dgamma), float(drow)
        self.nw, self.nh = int(nw), int(nh)
        self.col_c = (nw - 1) / 2.0 if col_c is None else float(col_c)
        self.row_c = (nh - 1) / 2.0 if row_c is None else float(row_c)

    @property
    def lambdas(self):
        return self.l0 + self.dl * np.arange(self.nviews)

    @property
    def src_z(self):
        return self.zs0 + self.hp * self.dl * np.arange(self.nviews)

    def geometry(self, vol_shape, vol_dim=None, vol_start=(0, 0, 0)):
        """Projector geometry (cylindrical records) for a volume block."""
        return Geometry.cylinder(self.lambdas, self.src_z, self.R, self.SDD, self.dgamma,
                                 self.drow, vol_shape, (self.nh, self.nw), col_c=self.col_c,
                                 row_c=self.row_c, vol_dim=vol_dim, vol_start=vol_start)

    @classmethod
    def config2(cls):
        """Workload configuration 2: voxel 0.5 mm; source radius 570 mm, source-to-detector
        1040 mm; 512 columns x 64 rows, column pitch 0.8 mm (arc) and row pitch 1.0 mm at the
        detector; 3 turns x 1440 views; pitch 0.8 x detector height at the isocentre. The helix
        is centred on the volume in z (the configuration leaves its placement open)."""
        vox = 0.5
        R, SDD = 570.0 / vox, 1040.0 / vox
        nw, nh = 512, 64
        dgamma = 0.8 / 1040.0
        drow = 1.0 / vox
        height_iso = nh * 1.0 * 570.0 / 1040.0 / vox       # voxels
        pitch = 0.8 * height_iso                             # voxels per turn
        hp = pitch / (2 * math.pi)
        views_per_turn, turns = 1440, 3
        n = views_per_turn * turns
        dl = 2 * math.pi / views_per_turn
        l0 = 0.0
        zs0 = -hp * dl * (n - 1) / 2.0
        return cls(n, dl, l0, zs0, hp, R, SDD, dgamma, drow, nw, nh)


def katsevich(proj, scan, vol_shape, origin=None, out=None, npsi=None, return_filtered=False):
    """Reconstruct [nz, ny, nx] (float32, accumulated into `out`) from helical line integrals
    proj [nviews, nh, nw] (float32 / bfloat16 / float16 on the device). origin = coordinate of
    voxel (0,0,0) as (x, y, z); default centres the block."""
    _need_device(proj, "proj")
    if tuple(proj.shape) != (scan.nviews, scan.nh, scan.nw):
        raise ValueError("proj shape does not match the scan")
    nz, ny, nx = vol_shape
    if origin is None:
        origin = (-(nx // 2), -(ny // 2), -(nz // 2))
    if out is None:
        out = torch.zeros(vol_shape, dtype=torch.float32, device=proj.device)
    _need_device(out, "out")
    npsi = 2 * scan.nh + 1 if npsi is None else int(npsi)
    g5 = torch.empty(proj.shape, dtype=torch.float32, device=proj.device) if return_filtered \
        else None
    L = _native.lib()
    rc = L.fc_katsevich_dev(proj.device.index, proj.data_ptr(), _dtype_code(proj, "proj"),
                            scan.nw, scan.nh, scan.nviews, scan.l0, scan.dl, scan.zs0, scan.hp,
                            scan.R, scan.SDD, scan.dgamma, scan.drow, scan.col_c, scan.row_c,
                            npsi, out.data_ptr(), _native.int_array([nx, ny, nz]),
                            _native.dbl_array(origin),
                            g5.data_ptr() if g5 is not None else None, _stream(proj.device))
    _native.check(rc)
    return (out, g5) if return_filtered else out
