"""Headline benchmark: cone-beam forward projection + FDK backprojection, 512^3 x 720 views.

Workload (configuration 1 of the workload set): 512^3 float32 volume (voxel 0.5 mm) holding a
seeded phantom of 40 random ellipsoids (semi-axes U[3,80] mm, attenuation U[0.002,0.03]/mm,
centres within 60 mm) plus N(0, 0.001^2)/mm texture; 720 views over 360 deg; flat detector
768 x 768, pitch 0.75 voxel at the axis; source-to-axis 1200 voxels (600 mm); ray step 0.5 voxel.

One step = forward projection (stage + ray-march kernel) -> FDK filtering (cosine weight +
ramp, one FFT kernel) -> FDK backprojection (stage + voxel-driven kernel), all device-resident.
value = W / step time, W = 512^3 * 720 voxel-view updates (Gvoxel-view/s).

Extra fields: per-kernel rates, the roofline of the dominant kernel (the projector: its LDS gather
bytes against the LDS bandwidth, with its inner loop's own ceilings and the HBM view beside), the
BF16 / FP16 storage variants, a CPU baseline (the FP64 oracle on the host cores, a bounded sample
of the same workload, extrapolated) with parity of the GPU result on that sample, and the CGLS
time-to-residual leg of the learned warm start.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-variants]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxel-proj updates/s fwd+backproj 512³×720; CGLS time-to-res 1e-3 vs plain"
N, NV, NDET, D, PW, H = 512, 720, 768, 1200.0, 0.75, 0.5
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip BF16/FP16 storage runs")
    ap.add_argument("--cpu-views", type=int, default=24)
    ap.add_argument("--cpu-slices", type=int, default=16)
    ap.add_argument("--views", type=int, default=NV,
                    help="views of the workload (default 720; fewer only for profiling runs)")
    ap.add_argument("--random-volume", action="store_true",
                    help="uniform random volume instead of the phantom (profiling runs)")
    ap.add_argument("--cgls-model", default=os.path.join(ROOT, "fastcone_amd", "models",
                                                         "unet_warmstart.npz"),
                    help="trained warm-start network for the CGLS time-to-residual leg")
    ap.add_argument("--no-cgls", action="store_true", help="skip the CGLS leg")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run a plain f32 copy (convert_kernel) of this many MiB, 3 times: "
                         "the known-size reference for HBM byte counters")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the projector / backprojector runs on exactly one GPU; "
                                   "sharding over several GPUs (views or z-slabs per GPU) is not "
                                   "implemented, --gpus must be 1"}))
        sys.exit(2)
    import torch
    from fastcone_amd import _native, ops, phantom
    from fastcone_amd.geometry import Geometry

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calib_mib > 0:
        src = torch.rand(args.copy_calib_mib << 18, device=dev)
        for _ in range(3):
            ops.convert(src, torch.float32)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calib_mib} MiB read + written per launch, 3 launches")
        del src
    t_setup = time.time()
    nv = args.views
    if args.random_volume:
        g0 = torch.Generator(device=dev)
        g0.manual_seed(0)
        vol = torch.rand((N, N, N), generator=g0, device=dev) * 0.02
    else:
        vol, _ = phantom.config_phantom(1, device=dev)
    ang = 2 * np.pi * np.arange(nv) / nv
    geom = Geometry.circular(ang, D, PW, (N, N, N), (NDET, NDET))
    norm = torch.full((nv,), 1.0 / nv, dtype=torch.float32, device=dev)
    proj = torch.zeros(geom.proj_shape, dtype=torch.float32, device=dev)
    rec = torch.zeros(geom.vol_shape, dtype=torch.float32, device=dev)
    samples = int(ops.sample_count(geom, H, dev).sum().item())
    W = float(N) ** 3 * nv
    log(f"setup {time.time() - t_setup:.1f}s; FP samples {samples:.4e}")

    def step(v, store=torch.float32):
        proj.zero_()
        ops.project(v, geom, step=H, out=proj)
        q = ops.fdk_filter(proj, geom, D, PW, out_dtype=store)
        rec.zero_()
        ops.backproject(q, geom, _native.FC_BP_FDK, out=rec, norm=norm)

    def timed(v, store, steps, warmup):
        for _ in range(warmup):
            step(v, store)
        torch.cuda.synchronize()
        _native.profile_enable(True)
        _native.profile_reset()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record()
        for _ in range(steps):
            step(v, store)
        ev1.record()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        ev_ms = ev0.elapsed_time(ev1) / steps
        fp_n, fp_ms = _native.profile_read("project")
        bp_n, bp_ms = _native.profile_read("backproject")
        fl_n, fl_ms = _native.profile_read("fdk_filter")
        _native.profile_enable(False)
        return wall, ev_ms, fp_ms / max(fp_n, 1), bp_ms / max(bp_n, 1), fl_ms / max(fl_n, 1)

    wall, ev_ms, fp_ms, bp_ms, flt_ms = timed(vol, torch.float32, args.steps, args.warmup)
    log(f"f32: step {wall * 1e3:.1f} ms (events {ev_ms:.1f}), FP kernel {fp_ms:.1f} ms, "
        f"filter {flt_ms:.1f} ms, BP kernel {bp_ms:.1f} ms")

    # Roofline of the dominant kernel (the projector). Its algorithmic traffic is the gather from
    # the LDS bricks: 8 taps x 4 B = 32 B per ray sample; the ceiling is the LDS bandwidth from the
    # published figures (128 B per clock per CU, 256 CUs, 2.4 GHz). The kernels' own inner loops
    # on resident bricks / windows (ray-coherent and random positions) and the HBM view
    # (compulsory bytes against 8 TB/s, plus the measured FETCH/WRITE bytes) are reported beside.
    peak_fp = _native.gather_bench(4, _native.FC_F32)
    peak_fp_random = _native.gather_bench(3, _native.FC_F32)
    peak_bp = _native.gather_bench(5, _native.FC_F32)         # ray-coherent BP inner loop
    peak_bp_random = _native.gather_bench(2, _native.FC_F32)
    lds_bw = 128.0 * 256 * 2.4e9                               # bytes / s
    fp_taps = 8.0 * samples
    bp_taps = 4.0 * W
    fp_rate = fp_taps / (fp_ms * 1e-3)
    # compulsory HBM bytes of one projector launch: read the staged (padded) volume once, read
    # and write the projection stack once
    fp_hbm_bytes = (N + 4) ** 3 * 4 + 2 * 4 * nv * NDET * NDET
    traffic = measured_traffic(nv)
    roofline = {
        "bound": "lds",
        "kernel": "fp_quad_kernel + fp_lds_kernel (ray-march projector, LDS slab bricks shared by "
                  "four-view quads)",
        "achieved": fp_rate * 4.0 / 1e9, "peak": lds_bw / 1e9, "unit": "GB/s",
        "frac": fp_rate * 4.0 / lds_bw,
        "traffic": None if traffic is None else traffic["bytes_per_launch"],
        "per_unit": "32 B of LDS gather per ray sample (8 trilinear taps x 4 B); samples = exact "
                    "count of the clipped rays",
        "units_per_launch": samples, "taps_per_s": fp_rate,
        "peak_kind": "LDS bandwidth: 128 B/clock/CU x 256 CUs x 2.4 GHz (published figures)",
        "limiter": "VALU issue (1.25-1.30 wave instructions per CU clock by the SQ counters in "
                   "profiles/fp_quad_kernel_pmc.md); a lane rereads its cell's corner pairs only "
                   "when the cell changes, so the LDS moves about 60 % of the algorithmic bytes",
        "ceilings": {
            "inner_loop_ray_coherent_GBps": peak_fp * 32 / 1e9,
            "inner_loop_ray_coherent_frac": fp_rate / (peak_fp * 8),
            "inner_loop_random_positions_GBps": peak_fp_random * 32 / 1e9},
        "hbm": {"achieved": fp_hbm_bytes / (fp_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": fp_hbm_bytes / (fp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "bytes_per_launch": fp_hbm_bytes,
                "measured": traffic},
    }
    kernels = {
        "fp": {"ms": fp_ms, "Gvoxel_view_per_s": W / (fp_ms * 1e-3) / 1e9,
               "Gtap_per_s": fp_rate / 1e9},
        "bp": {"ms": bp_ms, "Gvoxel_view_per_s": W / (bp_ms * 1e-3) / 1e9,
               "Gtap_per_s": bp_taps / (bp_ms * 1e-3) / 1e9,
               "lds_GBps": bp_taps * 4.0 / (bp_ms * 1e-3) / 1e9,
               "lds_frac": bp_taps * 4.0 / (bp_ms * 1e-3) / lds_bw,
               "gather_peak_Gtap_per_s": peak_bp * 4 / 1e9,
               "gather_frac": bp_taps / (bp_ms * 1e-3) / (peak_bp * 4),
               "gather_random_positions_Gtap_per_s": peak_bp_random * 4 / 1e9},
        "filter": {"ms": flt_ms, "kernel": "row_filter_kernel (cosine weight + ramp, FFT in LDS)",
                   "hbm_GBps": 2 * 4 * nv * NDET * NDET / (flt_ms * 1e-3) / 1e9 if flt_ms else None},
        "R_kernels_only": W / ((fp_ms + bp_ms) * 1e-3) / 1e9,
    }

    variants = {}
    if not args.no_variants:
        for name, dt in (("bf16", torch.bfloat16), ("f16", torch.float16)):
            v16 = vol.to(dt)
            w16, e16, f16ms, b16ms, _ = timed(v16, dt, max(1, args.steps - 1), 1)
            variants[name] = {"ms_per_step": w16 * 1e3, "value": W / w16 / 1e9,
                              "fp_ms": f16ms, "bp_ms": b16ms,
                              "fp_gather_peak_Gtap_per_s": _native.gather_bench(
                                  0, _native.FC_BF16 if dt == torch.bfloat16 else _native.FC_F16)
                              * 8 / 1e9}
            del v16
            log(f"{name}: step {w16 * 1e3:.1f} ms, FP {f16ms:.1f} ms, BP {b16ms:.1f} ms")

    cpu = None
    parity = None
    if not args.no_cpu and nv == NV:
        cpu, parity = cpu_baseline(args, vol, geom, proj, rec, samples)

    cgls_leg = None
    if not args.no_cgls and nv == NV and os.path.exists(args.cgls_model):
        del proj, rec
        torch.cuda.empty_cache()
        from fastcone_amd import warmstart
        from fastcone_amd.unet import UNet3D
        net = UNet3D().to(dev).load(args.cgls_model)
        cgls_leg = warmstart.evaluate(net, dev, log=log)

    out = {
        "metric": METRIC, "value": W / wall / 1e9, "unit": "Gvoxel-view/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom)",
        "config": {"workload": "circular cone-beam 512^3 x 720 views, 768^2 flat detector, "
                               "FP step 0.5 voxel + FDK backprojection, FP32 storage",
                   "volume": [N, N, N], "views": NV, "detector": [NDET, NDET],
                   "pixel_pitch_vox": PW, "source_axis_vox": D, "step_vox": H,
                   "W_voxel_views": W, "fp_samples": samples},
        "kernels": kernels, "roofline": roofline, "variants": variants,
        "cpu_baseline": cpu, "parity_on_cpu_sample": parity, "cgls_time_to_residual": cgls_leg,
    }
    print(json.dumps(out), flush=True)


def measured_traffic(nv):
    """HBM bytes per projector launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (each scaled by what the counter reads for a copy kernel of known size), when they
    were measured at this view count; None otherwise."""
    path = os.path.join(ROOT, "profiles", "fp_quad_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if int(t.get("views", -1)) != nv:
        return None
    return t


def cpu_baseline(args, vol, geom, proj, rec, samples):
    """FP64 OpenMP oracle on the host cores: FP of a few views (full detector) and FDK BP of a few
    z-slices (all views), timed, extrapolated to the whole workload; also the relative L2 of the
    GPU result (last FP32 step) on that sample."""
    import oracle
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    oracle.build()
    threads = oracle.num_threads()
    nv, ns = args.cpu_views, args.cpu_slices
    idx = np.linspace(0, NV - 1, nv).astype(int)
    ang = 2 * np.pi * np.arange(NV) / NV
    vol_np = vol.cpu().numpy().astype(np.float64)
    prob_fp = oracle.Problem(ang[idx], np.full(nv, N / 2), D, PW, vol_dim=(N, N, N),
                             proj_dim=(NDET, NDET))
    t0 = time.perf_counter()
    p_ref = oracle.project(prob_fp, vol_np, h=H)
    t_fp = time.perf_counter() - t0
    # BP sample: ns z-slices through the centre, all views, of the filtered stack of this run
    from fastcone_amd import ops
    q = ops.fdk_filter(proj, geom, D, PW).cpu().numpy().astype(np.float64)
    z0 = N // 2 - ns // 2
    prob_bp = oracle.Problem(ang, np.full(NV, N / 2), D, PW, vol_dim=(N, N, N),
                             vol_start=(0, 0, z0), proj_dim=(NDET, NDET))
    t0 = time.perf_counter()
    v_ref = oracle.backproject(prob_bp, q, (ns, N, N), mode=2, norm=np.full(NV, 1.0 / NV))
    t_bp = time.perf_counter() - t0
    t_fp_full = t_fp * NV / nv
    t_bp_full = t_bp * N / ns
    W = float(N) ** 3 * NV
    p_gpu = proj.cpu().numpy()[idx].astype(np.float64)
    v_gpu = rec.cpu().numpy()[z0:z0 + ns].astype(np.float64)
    parity = {"fp_rel_l2": float(np.linalg.norm(p_gpu - p_ref) / np.linalg.norm(p_ref)),
              "bp_rel_l2": float(np.linalg.norm(v_gpu - v_ref) / np.linalg.norm(v_ref)),
              "fp_views": [int(i) for i in idx], "bp_slices": [int(z0), int(z0 + ns)]}
    cpu = {"value": W / (t_fp_full + t_bp_full) / 1e9, "unit": "Gvoxel-view/s",
           "cores": threads, "kind": "port",
           "sample": f"FP64 oracle: forward projection of {nv} of {NV} views (full 768^2 "
                     f"detector, {t_fp:.1f} s) and FDK backprojection of {ns} of {N} z-slices "
                     f"with all {NV} views ({t_bp:.1f} s), each scaled linearly to the full "
                     f"workload",
           "fp_s_sample": t_fp, "bp_s_sample": t_bp}
    log(f"cpu baseline: {cpu['value']:.4f} Gvoxel-view/s on {threads} threads; parity {parity}")
    return cpu, parity


if __name__ == "__main__":
    main()
