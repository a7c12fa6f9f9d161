"""CGLS for  min_x |A x - b|^2  with the matched projector pair, FP64 inner products on the device.

The reference has no iterative solver; this is the Hestenes-Stiefel CGLS recursion
(Bjorck, "Numerical Methods for Least Squares Problems", alg. 7.4.1):

    r = b - A x,  s = A^T r,  p = s,  gamma = |s|^2
    repeat:  q = A p;  alpha = gamma / |q|^2;  x += alpha p;  r -= alpha q
             s = A^T r;  gamma' = |s|^2;  p = s + (gamma'/gamma) p

Vectors stay float32 on the device (the operators' storage); every inner product and norm is
accumulated in float64 on the device. The stopping rule is the normal-equation residual
|A^T (b - A x_k)| / |A^T b| <= tol, whose denominator does not depend on the starting point.
"""
import time

import numpy as np
import torch


def dot64(a, b):
    """<a, b> accumulated in float64 on a's device."""
    return torch.dot(a.reshape(-1).double(), b.reshape(-1).double())


def norm64(a):
    return torch.sqrt(dot64(a, a))


class CGLSState:
    """Iteration state (x, r, p, gamma); save()/load() checkpoint it to .npz."""

    def __init__(self, x, r, p, gamma, it=0, atb_norm=None):
        self.x, self.r, self.p = x, r, p
        self.gamma = gamma
        self.it = it
        self.atb_norm = atb_norm

    def save(self, path):
        np.savez(path, x=self.x.cpu().numpy(), r=self.r.cpu().numpy(), p=self.p.cpu().numpy(),
                 gamma=float(self.gamma), it=self.it,
                 atb_norm=float(self.atb_norm) if self.atb_norm is not None else np.nan)

    @classmethod
    def load(cls, path, device):
        z = np.load(path)
        atb = float(z["atb_norm"])
        return cls(torch.from_numpy(z["x"]).to(device), torch.from_numpy(z["r"]).to(device),
                   torch.from_numpy(z["p"]).to(device), torch.tensor(float(z["gamma"]),
                                                                     dtype=torch.float64,
                                                                     device=device),
                   int(z["it"]), None if np.isnan(atb) else atb)


def _finite(v, what):
    if not bool(torch.isfinite(v)):
        raise FloatingPointError(f"CGLS: {what} is not finite (non-finite data or start point)")


def cgls(A, At, b, x0=None, iters=50, tol=None, atb_norm=None, state=None, callback=None,
         sync_every=1):
    """Run CGLS. A(x) -> projections, At(r) -> volume (exact transpose of A).

    Returns (x, info) with info = {"residual": [|A^T r_k| / |A^T b|], "rnorm": [|r_k|],
    "iters", "time_s", "converged", "state"}. `state` resumes a checkpointed run. Raises
    FloatingPointError when an inner product turns non-finite (NaN or infinite data).
    """
    t0 = time.perf_counter()
    if state is None:
        if x0 is None:
            r = b.clone()
            s = At(r)
            x = torch.zeros_like(s)
        else:
            x = x0.clone()
            r = b - A(x)
            s = At(r)
        p = s.clone()
        gamma = dot64(s, s)
        _finite(gamma, "|A^T r|^2")
        if atb_norm is None:
            atb_norm = float(torch.sqrt(gamma)) if x0 is None else float(norm64(At(b)))
        it0 = 0
    else:
        x, r, p, gamma, it0 = state.x, state.r, state.p, state.gamma, state.it
        atb_norm = state.atb_norm if atb_norm is None else atb_norm
    hist = [float(torch.sqrt(gamma)) / atb_norm]
    rn = [float(norm64(r))]
    converged = tol is not None and hist[-1] <= tol
    it = it0
    while not converged and it < iters:
        q = A(p)
        qq = dot64(q, q)
        _finite(qq, "|A p|^2")
        alpha = (gamma / qq).float()
        x.add_(alpha * p)
        r.sub_(alpha * q)
        s = At(r)
        gamma_new = dot64(s, s)
        _finite(gamma_new, "|A^T r|^2")
        beta = (gamma_new / gamma).float()
        p.mul_(beta).add_(s)
        gamma = gamma_new
        it += 1
        res = float(torch.sqrt(gamma)) / atb_norm
        hist.append(res)
        rn.append(float(norm64(r)))
        if callback is not None:
            callback(it, x, res)
        if tol is not None and res <= tol:
            converged = True
    info = {"residual": hist, "rnorm": rn, "iters": it - it0,
            "time_s": time.perf_counter() - t0, "converged": converged,
            "state": CGLSState(x, r, p, gamma, it, atb_norm)}
    return x, info


def poisson_log(proj, I0, scale=1.0, seed=0):
    """Noisy line integrals: counts ~ Poisson(I0 exp(-scale * p)), p_hat = -ln(max(counts,1)/I0)
    / scale (scale converts the projector's voxel-unit integrals to the attenuation units)."""
    g = torch.Generator(device=proj.device)
    g.manual_seed(int(seed))
    lam = I0 * torch.exp(-scale * proj.double())
    counts = torch.poisson(lam, generator=g)
    return (-torch.log(torch.clamp(counts, min=1.0) / I0) / scale).float()
