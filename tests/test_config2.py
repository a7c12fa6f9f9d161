"""Configuration 2 at full size: helical cone beam, 512 x 512 x 256 volume (0.5 mm voxels),
curved detector 512 x 64, 3 turns x 1440 views, pitch 0.8 x detector height at the isocentre.

The seeded phantom is projected by libFastcone's helical projector and reconstructed by
fc_katsevich_dev (derivative, kappa-line Hilbert filtering, PI-window backprojection). The FP64
oracle reconstructs two central z-slices from the same projections (all views); the GPU
reconstruction must match it to 1e-3 relative L2 (FP32 filtering against FP64), and both are
compared with the phantom inside the region the helix reconstructs exactly.
"""
import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_config2_full_size(gpu):
    from fastcone_amd import ops, phantom
    from fastcone_amd.katsevich import HelixScan, katsevich
    scan = HelixScan.config2()
    vol, ell = phantom.config_phantom(2, device=gpu)         # [256, 512, 512]
    nz, ny, nx = vol.shape
    origin = (-(nx // 2), -(ny // 2), -(nz // 2))
    g = scan.geometry(tuple(vol.shape))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    proj = ops.project(vol, g, step=0.5)
    torch.cuda.synchronize()
    t_fp = time.perf_counter() - t0
    katsevich(proj, scan, tuple(vol.shape), origin=origin)   # warm-up (FFT-free, but workspaces)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec = katsevich(proj, scan, tuple(vol.shape), origin=origin)
    torch.cuda.synchronize()
    t_kat = time.perf_counter() - t0
    # FP64 oracle on two central slices, all views, from the same projections
    z0 = nz // 2 - 1
    hx = oracle.Helix(scan.lambdas, scan.src_z, scan.R, scan.SDD, scan.dgamma, scan.drow,
                      scan.nw, scan.nh, scan.hp, scan.col_c, scan.row_c)
    p64 = proj.cpu().numpy().astype(np.float64)
    t0 = time.perf_counter()
    ref = oracle.katsevich(hx, p64, (2, ny, nx), np.array([origin[0], origin[1], origin[2] + z0],
                                                            dtype=np.float64))
    t_cpu = time.perf_counter() - t0
    got = rec[z0:z0 + 2].cpu().numpy().astype(np.float64)
    e_rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    # the ellipsoids without the per-voxel texture (white noise the chain low-passes)
    smooth = phantom.voxelize(ell, tuple(vol.shape), 0.5, gpu)
    truth = smooth[z0:z0 + 2].cpu().numpy().astype(np.float64)
    yy, xx = np.meshgrid(np.arange(ny) - ny // 2, np.arange(nx) - nx // 2, indexing="ij")
    fov = np.broadcast_to(np.hypot(xx, yy) < 0.9 * scan.R * np.sin(scan.dgamma * scan.nw / 2),
                          got.shape)
    e_ph = float(np.abs(got - truth)[fov].mean() / np.abs(truth[fov]).mean())
    print(f"config 2: helical FP {t_fp * 1e3:.0f} ms, Katsevich {t_kat * 1e3:.0f} ms on the GPU; "
          f"FP64 oracle (2 slices, {oracle.num_threads()} threads) {t_cpu:.1f} s; GPU vs oracle "
          f"rel L2 {e_rel:.2e}; mean |rec - phantom| / mean |phantom| in the field of view "
          f"{e_ph:.3f}")
    assert e_rel < 1e-3
    assert e_ph < 0.1
