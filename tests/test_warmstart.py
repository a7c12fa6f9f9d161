"""Learned warm start for CGLS (configuration 4): the committed network, trained by
`python -m fastcone_amd.warmstart train` on FDK -> 200-iteration-CGLS pairs at 128^3 / 90 views,
against plain FDK-start CGLS at 256^3 / 120 views, both run to normal-equation residual 1e-3.
The test reports what the network gains (wall-time ratio, start residuals) and checks the
fallback rule: a warm start whose residual exceeds the FDK start's is replaced by plain CGLS."""
import os

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MODEL = os.path.join(os.path.dirname(__file__), os.pardir, "fastcone_amd", "models",
                     "unet_warmstart.npz")


def test_warm_start_against_plain_cgls(gpu):
    from fastcone_amd import warmstart
    from fastcone_amd.unet import UNet3D
    net = UNet3D().to(gpu).load(MODEL)
    res = warmstart.evaluate(net, gpu, seed=777, tol=1e-3, log=lambda s: None)
    print(f"warm start: plain {res['plain_s']:.2f} s ({res['plain_iters']} iterations), warm "
          f"{res['warm_s']:.2f} s ({res['warm_iters']} iterations), speedup {res['speedup']:.2f}; "
          f"start residual FDK {res['fdk_start_residual']:.4f}, network "
          f"{res['net_start_residual']:.4f}, fallback {res['fallback']}")
    assert res["plain_converged"] and res["warm_converged"]
    assert res["fallback"] == (res["net_start_residual"] > res["fdk_start_residual"])
    if res["fallback"]:
        assert res["warm_iters"] == res["plain_iters"]
    else:
        assert res["warm_iters"] <= res["plain_iters"]
