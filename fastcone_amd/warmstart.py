"""Learned warm start for sparse-view CGLS (workload configuration 4).

Training pairs come from the method itself: a seeded random-ellipsoid phantom at 128^3 is
projected (90 views over 360 deg, 192^2 detector, geometry of configuration 0), Poisson noise
(I0 = 1e4 photons per ray) and the log transform are applied, and the pair is (FDK of the noisy
data, 200-iteration CGLS of the noisy data). The 3D U-Net (unet.py) learns FDK -> CGLS with an L2
loss (Adam, lr 1e-3, batch 4). At inference (256^3, 120 views, 384^2 detector) CGLS is started from
the network's output and from FDK and both are run to the normal-equation residual
|A^T (b - A x)| / |A^T b| <= 1e-3; if the warm start's initial residual exceeds the FDK start's,
plain FDK-start CGLS is used (fallback). Wall time of the warm start includes FDK and inference.

    python -m fastcone_amd.warmstart train --pairs 2000 --epochs 10 --out unet.npz
    python -m fastcone_amd.warmstart eval --model unet.npz

Training can be streamed over several runs: `--first-pair S` selects the phantom seeds
S .. S + pairs - 1 and `--resume CKPT` continues from the weights and Adam state that a previous
run saved next to its `--out` file (CKPT.npz + CKPT.adam.pt), so a long schedule of fresh pairs
can be split into runs of bounded length.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

from . import ops, phantom
from .cgls import cgls, norm64, poisson_log
from .geometry import Geometry
from .unet import UNet3D

I0 = 1e4
STEP = 0.5


def train_geometry():
    ang = 2 * np.pi * np.arange(90) / 90
    return Geometry.circular(ang, 500.0, 0.75, (128, 128, 128), (192, 192))


def infer_geometry():
    ang = 2 * np.pi * np.arange(120) / 120
    return Geometry.circular(ang, 500.0, 0.6, (256, 256, 256), (384, 384))


def operators(geom):
    A = lambda v: ops.project(v, geom, STEP)
    At = lambda r: ops.adjoint(r, geom, STEP)
    return A, At


def fdk(b, geom):
    pw = float(np.linalg.norm(geom.views[0, 6:9]))
    D = float(np.linalg.norm(geom.views[0, 0:2]))
    return ops.fdk(b, geom, D, pw)


def make_pair(seed, geom, device, cgls_iters=200):
    """(FDK input, CGLS target, phantom) for one seeded phantom, all FP32 on the device."""
    shape = geom.vol_shape
    ell = phantom.random_ellipsoids(10_000 + seed, 10, (5, 40), (0.005, 0.02), 40.0)
    vol = phantom.voxelize(ell, shape, 1.0, device)
    A, At = operators(geom)
    b = poisson_log(A(vol), I0, scale=1.0, seed=seed)
    x_fdk = fdk(b, geom)
    x_cg, _ = cgls(A, At, b, iters=cgls_iters)
    return x_fdk, x_cg, vol


def generate(npairs, device, cgls_iters=200, log=print, first=0, time_budget=None):
    """Pairs for phantom seeds first .. first + npairs - 1 (BF16 on the device); stops early once
    time_budget seconds are spent and returns the pairs made so far."""
    geom = train_geometry()
    n = geom.vol_shape
    X = torch.empty((npairs,) + n, dtype=torch.bfloat16, device=device)
    Y = torch.empty((npairs,) + n, dtype=torch.bfloat16, device=device)
    t0 = time.perf_counter()
    made = 0
    for s in range(npairs):
        xf, xc, _ = make_pair(first + s, geom, device, cgls_iters)
        X[s] = xf.to(torch.bfloat16)
        Y[s] = xc.to(torch.bfloat16)
        made = s + 1
        if s % 25 == 0 or s == npairs - 1:
            torch.cuda.synchronize()
            log(f"pairs {s + 1}/{npairs} ({time.perf_counter() - t0:.1f} s)")
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            torch.cuda.synchronize()
            log(f"generation time budget reached after {made} pairs")
            break
    return X[:made], Y[:made]


def train(X, Y, epochs, batch=4, lr=1e-3, seed=0, log=print, val_frac=0.05, resume=None):
    net = UNet3D(seed=seed).to(X.device)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    if resume:
        net.load(resume + ".npz")
        opt.load_state_dict(torch.load(resume + ".adam.pt", weights_only=True))
    n = X.shape[0]
    nval = max(1, int(round(n * val_frac)))
    tr = np.arange(n - nval)
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        rng.shuffle(tr)
        tot = 0.0
        for q in range(0, len(tr) - batch + 1, batch):
            idx = torch.from_numpy(tr[q:q + batch]).to(X.device)
            xb = X[idx].float()
            yb = Y[idx].float()
            pred = net(xb)
            loss = ((pred - yb) ** 2).mean() * UNet3D.SCALE ** 2
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            tot += float(loss.detach())
        with torch.no_grad():
            xv = X[n - nval:].float()
            yv = Y[n - nval:].float()
            pv = torch.cat([net(xv[k:k + batch]) for k in range(0, nval, batch)])
            val = float(((pv - yv) ** 2).mean() / ((xv - yv) ** 2).mean())
        hist.append({"epoch": ep, "train_loss": tot / max(1, len(tr) // batch),
                     "val_mse_vs_fdk_mse": val, "s": time.perf_counter() - t0})
        log(json.dumps(hist[-1]))
    return net, hist, opt


def time_to_residual(geom, b, x0, tol=1e-3, max_iters=400, atb_norm=None):
    A, At = operators(geom)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, info = cgls(A, At, b, x0=x0, iters=max_iters, tol=tol, atb_norm=atb_norm)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, info


def evaluate(net, device, seed=777, tol=1e-3, log=print):
    """CGLS from the network output vs from FDK at 256^3 / 120 views, both to tol."""
    geom = infer_geometry()
    A, At = operators(geom)
    ell = phantom.random_ellipsoids(seed, 10, (5, 40), (0.005, 0.02), 40.0)
    vol = phantom.voxelize(ell, geom.vol_shape, 1.0, device)
    b = poisson_log(A(vol), I0, scale=1.0, seed=seed)
    atb = float(norm64(At(b)))
    # untimed warm-up of every piece (FFT plans, workspaces, kernels)
    with torch.no_grad():
        net(fdk(b, geom))
    torch.cuda.synchronize()
    # plain: FDK start
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x_fdk = fdk(b, geom)
    torch.cuda.synchronize()
    t_fdk = time.perf_counter() - t0
    t_plain, info_plain = time_to_residual(geom, b, x_fdk, tol, atb_norm=atb)
    # warm: FDK + network
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        x_net = net(fdk(b, geom))
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    r_net = float(norm64(At(b - A(x_net)))) / atb
    r_fdk = info_plain["residual"][0]
    fallback = r_net > r_fdk
    if fallback:
        t_warm, info_warm = t_plain, info_plain
    else:
        t_warm, info_warm = time_to_residual(geom, b, x_net, tol, atb_norm=atb)
    total_plain = t_fdk + t_plain
    total_warm = t_inf + t_warm
    res = {"tol": tol, "plain_s": total_plain, "warm_s": total_warm,
           "speedup": total_plain / total_warm, "plain_iters": info_plain["iters"],
           "warm_iters": info_warm["iters"], "fdk_start_residual": r_fdk,
           "net_start_residual": r_net, "fallback": bool(fallback), "inference_s": t_inf,
           "fdk_s": t_fdk, "plain_converged": info_plain["converged"],
           "warm_converged": info_warm["converged"]}
    log(json.dumps(res))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval", "pair"])
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--cgls-iters", type=int, default=200)
    ap.add_argument("--model", default=None)
    ap.add_argument("--out", default="unet_warmstart.npz")
    ap.add_argument("--seed", type=int, default=777)
    ap.add_argument("--first-pair", type=int, default=0, help="first phantom seed of this run")
    ap.add_argument("--resume", default=None,
                    help="checkpoint prefix (PREFIX.npz + PREFIX.adam.pt) to continue from")
    ap.add_argument("--time-budget", type=float, default=None,
                    help="seconds of pair generation after which training starts anyway")
    ap.add_argument("--no-eval", action="store_true")
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    log = lambda s: print(s, flush=True)
    if a.cmd == "pair":
        t0 = time.perf_counter()
        xf, xc, vol = make_pair(0, train_geometry(), dev, a.cgls_iters)
        torch.cuda.synchronize()
        log(json.dumps({"pair_s": time.perf_counter() - t0,
                        "fdk_err": float((xf - vol).norm() / vol.norm()),
                        "cgls_err": float((xc - vol).norm() / vol.norm())}))
        return
    if a.cmd == "train":
        X, Y = generate(a.pairs, dev, a.cgls_iters, log, first=a.first_pair,
                        time_budget=a.time_budget)
        net, hist, opt = train(X, Y, a.epochs, log=log, resume=a.resume)
        prefix = a.out[:-4] if a.out.endswith(".npz") else a.out
        net.save(prefix + ".npz")
        torch.save(opt.state_dict(), prefix + ".adam.pt")
        log(json.dumps({"pairs": int(X.shape[0]), "first_pair": a.first_pair,
                        "resumed_from": a.resume}))
        del X, Y
        if not a.no_eval:
            evaluate(net, dev, a.seed, log=log)
        return
    net = UNet3D().to(dev)
    if a.model:
        net.load(a.model)
    evaluate(net, dev, a.seed, log=log)


if __name__ == "__main__":
    main()
