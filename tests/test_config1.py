"""Configuration 1 at full size: circular cone beam, 512^3 volume (0.5 mm voxels), 720 views over
360 deg, flat detector 768 x 768 (pitch 0.75 voxel at the axis), source-to-axis 1200 voxels, ray
step 0.5 voxel; the seeded 40-ellipsoid phantom with texture. Forward projection, FDK filtering
(cosine weight + ramp, the HIP FFT kernel) and FDK backprojection, with FP32, BF16 and FP16
storage (the volume for the projector, the filtered projections for the backprojector; FP32
accumulation everywhere), each against the FP64 oracle:

  * FP: six whole views (all 768 rows, so the top and bottom detector rows are included) against
    the oracle's projection of the unrounded phantom;
  * filter: the GPU filter of the GPU's FP32 projections against the oracle's physical cosine
    weight + ramp / pw of the same projections (the six views), with a library FP32 FFT
    (torch.fft) of the same rows reported beside it;
  * BP: z-slices {0, 1, 255, 256, 510, 511} (the largest cone angles and the mid-plane) against
    the oracle's FP64 filter + FP64 FDK backprojection of the GPU's FP32 projections, all views.

FP32 storage must be within 1e-4 (the workload's target); the 16-bit errors are printed per result
and held to bounds set by their storage rounding.
"""
import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N, NV, NDET, D, PW, H = 512, 720, 768, 1200.0, 0.75, 0.5
VIEWS = [0, 90, 181, 359, 450, 719]
SLABS = [(0, 2), (255, 2), (510, 2)]


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_config1_full_size_parity_all_storage_types(gpu):
    from fastcone_amd import _native, ops, phantom
    from fastcone_amd.geometry import Geometry
    t_start = time.perf_counter()
    vol, _ = phantom.config_phantom(1, device=gpu)
    ang = 2 * np.pi * np.arange(NV) / NV
    g = Geometry.circular(ang, D, PW, (N, N, N), (NDET, NDET))
    norm = torch.full((NV,), 1.0 / NV, device=gpu)
    sub = oracle.Problem(ang[VIEWS], np.full(len(VIEWS), N / 2), D, PW, vol_dim=(N, N, N),
                         proj_dim=(NDET, NDET))
    vol_np = vol.cpu().numpy()
    t0 = time.perf_counter()
    ref_fp = oracle.project(sub, vol_np, h=H)
    t_fp_cpu = time.perf_counter() - t0

    report = {}
    ref_bp = None
    for name in ("float32", "bfloat16", "float16"):
        dt = getattr(torch, name)
        v = vol if dt == torch.float32 else ops.convert(vol, dt)
        p = ops.project(v, g, step=H)
        p_s = p[VIEWS].cpu().numpy()
        r = {"fp": rel(p_s, ref_fp)}
        if dt != torch.float32:
            # the arithmetic alone: the oracle fed the same rounded volume
            r["fp_vs_rounded_input"] = rel(p_s, oracle.project(sub, v.float().cpu().numpy(), h=H))
        q = ops.fdk_filter(p, g, D, PW, out_dtype=dt)
        if dt == torch.float32:
            q_ref = oracle.fdk_filter(sub, p_s)
            r["filter"] = rel(q[VIEWS].cpu().numpy(), q_ref)
            # yardstick: a library FP32 FFT (torch.fft) of the same weighted rows, same ramp
            w = torch.from_numpy(oracle.fdk_cosine_weight(sub)).to(gpu)
            x = p[VIEWS] * w.float()
            Hr = torch.from_numpy(oracle.ramp_filter_response(NDET) / PW).to(gpu).float()
            qt = torch.fft.irfft(torch.fft.rfft(x, n=2 * NDET, dim=2) * Hr, n=2 * NDET,
                                 dim=2)[..., :NDET]
            r["filter_torch_fft_f32"] = rel(qt.cpu().numpy(), q_ref)
            del w, x, qt
            # FP64 reference chain for the backprojection: oracle filter + oracle FDK BP of the
            # GPU's FP32 projections (all views)
            t0 = time.perf_counter()
            q64 = oracle.fdk_filter(oracle.Problem(ang, np.full(NV, N / 2), D, PW,
                                                   vol_dim=(N, N, N), proj_dim=(NDET, NDET)),
                                    p.cpu().numpy())
            ref_bp = []
            for z0, nz in SLABS:
                prob = oracle.Problem(ang, np.full(NV, N / 2), D, PW, vol_dim=(N, N, N),
                                      vol_start=(0, 0, z0), proj_dim=(NDET, NDET))
                ref_bp.append(oracle.backproject(prob, q64, (nz, N, N), mode=2,
                                                 norm=np.full(NV, 1.0 / NV)))
            del q64
            t_bp_cpu = time.perf_counter() - t0
        del p
        rec = ops.backproject(q, g, _native.FC_BP_FDK, norm=norm)
        del q
        got = [rec[z0:z0 + nz].cpu().numpy() for z0, nz in SLABS]
        r["bp"] = rel(np.concatenate(got), np.concatenate(ref_bp))
        r["bp_per_slab"] = [rel(a, b) for a, b in zip(got, ref_bp)]
        del rec, v
        torch.cuda.empty_cache()
        report[name] = r
        print(f"config 1 {name}: " + ", ".join(
            f"{k} {val:.2e}" if isinstance(val, float) else f"{k} {['%.2e' % x for x in val]}"
            for k, val in r.items()), flush=True)
    print(f"config 1: oracle FP of {len(VIEWS)} views {t_fp_cpu:.1f} s, oracle filter + BP of "
          f"{sum(n for _, n in SLABS)} slices {t_bp_cpu:.1f} s on {oracle.num_threads()} threads; "
          f"total {time.perf_counter() - t_start:.1f} s")
    f32, b16, h16 = report["float32"], report["bfloat16"], report["float16"]
    # the filter's FP32 FFT rounding is relative to the input rows, whose energy sits at low
    # frequencies where the ramp is small: ~1e-5 of the (much smaller) output
    assert f32["fp"] < 1e-4 and f32["filter"] < 5e-5 and f32["bp"] < 1e-4
    assert max(f32["bp_per_slab"]) < 1e-4
    assert b16["fp_vs_rounded_input"] < 5e-5 and h16["fp_vs_rounded_input"] < 5e-5
    assert b16["fp"] < 5e-3 and b16["bp"] < 2e-2
    assert h16["fp"] < 1e-3 and h16["bp"] < 5e-3
