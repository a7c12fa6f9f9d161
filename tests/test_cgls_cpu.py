"""CGLS recursion on small dense least-squares problems (CPU, torch float32 vectors with FP64
reductions) against scipy's direct least-squares solution."""
import numpy as np
import pytest
import scipy.linalg
import torch

from fastcone_amd.cgls import CGLSState, cgls


def _problem(m=60, n=25, seed=0, cond=30.0):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, m)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sv = np.geomspace(1.0, 1.0 / cond, n)
    M = (U[:, :n] * sv) @ V.T
    b = rng.standard_normal(m)
    return M, b


def test_cgls_converges_to_least_squares_solution():
    M, b = _problem()
    Mt = torch.from_numpy(M.astype(np.float32))
    A = lambda x: Mt @ x
    At = lambda r: Mt.T @ r
    x, info = cgls(A, At, torch.from_numpy(b.astype(np.float32)), iters=200, tol=1e-5)
    xs = scipy.linalg.lstsq(M, b)[0]
    assert info["converged"]
    assert np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs) < 1e-3
    # |r_k| is non-increasing for CGLS
    rn = np.array(info["rnorm"])
    assert np.all(np.diff(rn) <= 1e-6 * rn[0])


def test_cgls_exact_in_n_steps_for_consistent_system():
    rng = np.random.default_rng(1)
    M = rng.standard_normal((12, 6))
    xt = rng.standard_normal(6)
    Mt = torch.from_numpy(M)
    x, info = cgls(lambda v: Mt @ v, lambda r: Mt.T @ r, torch.from_numpy(M @ xt), iters=6)
    assert np.allclose(x.numpy(), xt, atol=1e-8)


def test_cgls_warm_start_and_checkpoint_resume(tmp_path):
    M, b = _problem(seed=2)
    Mt = torch.from_numpy(M)
    A = lambda x: Mt @ x
    At = lambda r: Mt.T @ r
    bt = torch.from_numpy(b)
    x_full, info_full = cgls(A, At, bt, iters=20)
    x5, info5 = cgls(A, At, bt, iters=8)
    info5["state"].save(tmp_path / "s.npz")
    st = CGLSState.load(tmp_path / "s.npz", torch.device("cpu"))
    x_res, info_res = cgls(A, At, bt, iters=20, state=st)
    assert np.allclose(x_res.numpy(), x_full.numpy(), atol=1e-10)
    # warm start from the solution: residual already at the floor
    xs = torch.from_numpy(scipy.linalg.lstsq(M, b)[0])
    _, info_w = cgls(A, At, bt, x0=xs, iters=3, atb_norm=float(torch.linalg.norm(At(bt))))
    assert info_w["residual"][0] < 1e-10


def test_poisson_log_statistics():
    # counts ~ Poisson(lam), lam = I0 exp(-s p): -ln(N / I0) / s has mean p + 1 / (2 lam s) and
    # variance 1 / (lam s^2) to leading order in 1 / lam (delta method)
    from fastcone_amd.cgls import poisson_log
    I0, n = 1e4, 400_000
    for p, s in ((1.0, 1.0), (0.4, 2.5)):
        lam = I0 * np.exp(-s * p)
        q = poisson_log(torch.full((n,), p), I0, scale=s, seed=7).double().numpy()
        assert abs(q.mean() - (p + 1.0 / (2.0 * lam * s))) < 5.0 * np.sqrt(1.0 / (lam * s * s) / n)
        assert abs(q.var() * lam * s * s - 1.0) < 0.02


def test_poisson_log_zero_counts_and_seed():
    from fastcone_amd.cgls import poisson_log
    p = torch.tensor([60.0, 0.0, 1.0])
    q = poisson_log(p, 1e4, seed=3)
    assert q[0].item() == pytest.approx(np.log(1e4), rel=1e-6)  # zero counts clamp to one
    assert torch.equal(poisson_log(p, 1e4, seed=3), q)
    assert not torch.equal(poisson_log(torch.ones(1000), 1e4, seed=3),
                           poisson_log(torch.ones(1000), 1e4, seed=4))
