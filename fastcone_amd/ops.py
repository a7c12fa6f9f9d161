"""Device operators on torch tensors (ROCm), backed by the fc_* entry points of libFastcone.so.

* ``project(vol, geom, step)``        forward projection (ray marching, trilinear)
* ``adjoint(proj, geom, step)``       its exact transpose (matched backprojector)
* ``backproject(proj, geom, mode)``   voxel-driven FDK / Katsevich / unweighted backprojection
* ``ConeBeamProject``                 torch.autograd.Function: forward = project, backward =
                                      adjoint (checked by dot-product and finite-difference tests)
* ``fdk_filter(proj, geom, D, pw)``   cosine weight + ramp filter (one HIP FFT kernel)
* ``fdk(proj, geom, D, pw)``          fdk_filter + FDK backprojection in physical units
                                      (reconstructs the attenuation per voxel)

Tensors must live on a ROCm device; there is no CPU path. Volumes are [nz, ny, nx] and projection
stacks [nviews, nh, nw], C-contiguous. Work is enqueued on torch's current stream.
"""
import numpy as np
import torch

from . import _native
from .geometry import Geometry

_DT = {torch.float32: _native.FC_F32, torch.bfloat16: _native.FC_BF16,
       torch.float16: _native.FC_F16}


def _dtype_code(t, what):
    try:
        return _DT[t.dtype]
    except KeyError:
        raise TypeError(f"{what}: unsupported dtype {t.dtype} (float32, bfloat16, float16)")


def _need_device(t, what):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{what} must be a ROCm (cuda) tensor; libFastcone has no CPU path")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")


def _stream(device):
    return ctypes_void(torch.cuda.current_stream(device).cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(int(v))


def _norm_ptr(norm, geom, device):
    if norm is None:
        return None, None
    if not isinstance(norm, torch.Tensor):
        norm = torch.as_tensor(np.asarray(norm, np.float32))
    norm = norm.to(device=device, dtype=torch.float32).contiguous()
    if norm.numel() != geom.nviews:
        raise ValueError("norm must have one entry per view")
    return norm, norm.data_ptr()


def _vol3(shape):
    nz, ny, nx = shape
    return _native.int_array([nx, ny, nz])


def _proj3(geom):
    nh, nw = geom.det_shape
    return _native.int_array([nw, nh, geom.nviews])


def project(vol, geom, step=1.0, out=None, norm=None):
    """proj[t,i,j] (+)= norm[t] * step * sum_k T(src_t + k*step*u_tij). Returns proj (float32)."""
    _need_device(vol, "vol")
    if tuple(vol.shape) != geom.vol_shape:
        raise ValueError(f"vol shape {tuple(vol.shape)} != geometry {geom.vol_shape}")
    dt = _dtype_code(vol, "vol")
    if out is None:
        out = torch.zeros(geom.proj_shape, dtype=torch.float32, device=vol.device)
    else:
        _need_device(out, "out")
        if out.dtype != torch.float32 or tuple(out.shape) != geom.proj_shape:
            raise ValueError("out must be float32 with the geometry's projection shape")
    norm_t, norm_p = _norm_ptr(norm, geom, vol.device)
    views = geom.views_on(vol.device)
    L = _native.lib()
    rc = L.fc_project_dev(vol.device.index, vol.data_ptr(), dt, _vol3(geom.vol_shape),
                          _native.dbl_array(geom.origin), out.data_ptr(), _proj3(geom),
                          views.data_ptr(), float(step), norm_p, _stream(vol.device))
    _native.check(rc)
    return out


def adjoint(proj, geom, step=1.0, out=None, norm=None):
    """vol (+)= A^T proj, the exact transpose of project(). Returns vol (float32)."""
    _need_device(proj, "proj")
    if tuple(proj.shape) != geom.proj_shape:
        raise ValueError(f"proj shape {tuple(proj.shape)} != geometry {geom.proj_shape}")
    dt = _dtype_code(proj, "proj")
    if out is None:
        out = torch.zeros(geom.vol_shape, dtype=torch.float32, device=proj.device)
    else:
        _need_device(out, "out")
        if out.dtype != torch.float32 or tuple(out.shape) != geom.vol_shape:
            raise ValueError("out must be float32 with the geometry's volume shape")
    norm_t, norm_p = _norm_ptr(norm, geom, proj.device)
    views = geom.views_on(proj.device)
    L = _native.lib()
    rc = L.fc_adjoint_dev(proj.device.index, out.data_ptr(), _vol3(geom.vol_shape),
                          _native.dbl_array(geom.origin), proj.data_ptr(), dt, _proj3(geom),
                          views.data_ptr(), float(step), norm_p, _stream(proj.device))
    _native.check(rc)
    return out


def backproject(proj, geom, mode=_native.FC_BP_FDK, out=None, norm=None):
    """vol (+)= sum_t norm_t w_t B_t(.) with w = 1, U or U^2 (mode 0, 1, 2). Returns vol."""
    _need_device(proj, "proj")
    if tuple(proj.shape) != geom.proj_shape:
        raise ValueError(f"proj shape {tuple(proj.shape)} != geometry {geom.proj_shape}")
    if geom.kind != "flat":
        raise ValueError("backproject() needs a flat-detector geometry (see katsevich module)")
    dt = _dtype_code(proj, "proj")
    if out is None:
        out = torch.zeros(geom.vol_shape, dtype=torch.float32, device=proj.device)
    else:
        _need_device(out, "out")
        if out.dtype != torch.float32 or tuple(out.shape) != geom.vol_shape:
            raise ValueError("out must be float32 with the geometry's volume shape")
    norm_t, norm_p = _norm_ptr(norm, geom, proj.device)
    views = geom.views_on(proj.device)
    L = _native.lib()
    rc = L.fc_backproject_dev(proj.device.index, out.data_ptr(), _vol3(geom.vol_shape),
                              _native.dbl_array(geom.origin), proj.data_ptr(), dt, _proj3(geom),
                              views.data_ptr(), norm_p, int(mode), _stream(proj.device))
    _native.check(rc)
    return out


def sample_count(geom, step, device=None):
    """Exact number of projector samples per view (int64 tensor on the device)."""
    device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    counts = torch.zeros(geom.nviews, dtype=torch.int64, device=device)
    views = geom.views_on(device)
    L = _native.lib()
    rc = L.fc_sample_count_dev(device.index, _vol3(geom.vol_shape),
                               _native.dbl_array(geom.origin), _proj3(geom), views.data_ptr(),
                               float(step), counts.data_ptr(), _stream(device))
    _native.check(rc)
    return counts


def convert(t, dtype):
    """Storage conversion on the device with round-to-nearest-even (libFastcone kernel)."""
    _need_device(t, "tensor")
    out = torch.empty(t.shape, dtype=dtype, device=t.device)
    L = _native.lib()
    rc = L.fc_convert_dev(t.device.index, t.data_ptr(), _dtype_code(t, "src"), out.data_ptr(),
                          _dtype_code(out, "dst"), t.numel(), _stream(t.device))
    _native.check(rc)
    return out


class ConeBeamProject(torch.autograd.Function):
    """y = A x with A the step-h ray-marching projector; dL/dx = A^T dL/dy (matched adjoint)."""

    @staticmethod
    def forward(ctx, vol, geom, step):
        ctx.geom = geom
        ctx.step = float(step)
        return project(vol.contiguous(), geom, step)

    @staticmethod
    def backward(ctx, grad):
        g = adjoint(grad.contiguous().float(), ctx.geom, ctx.step)
        return g, None, None


class ConeBeamAdjoint(torch.autograd.Function):
    """x = A^T y; its gradient is the projector."""

    @staticmethod
    def forward(ctx, proj, geom, step):
        ctx.geom = geom
        ctx.step = float(step)
        return adjoint(proj.contiguous(), geom, step)

    @staticmethod
    def backward(ctx, grad):
        return project(grad.contiguous().float(), ctx.geom, ctx.step), None, None


def cone_project(vol, geom, step=1.0):
    return ConeBeamProject.apply(vol, geom, step)


def cone_adjoint(proj, geom, step=1.0):
    return ConeBeamAdjoint.apply(proj, geom, step)


# ---- FDK ------------------------------------------------------------------------------------
def fdk_filter(proj, geom, D=None, pw=1.0, out_dtype=torch.float32, out=None, cosine=True):
    """Cosine pre-weighting and ramp filtering of a projection stack on the device (one HIP kernel,
    fc_fdk_filter_dev: FFT convolution of every detector row in LDS, FP32 arithmetic).

    The ramp is the reference's (fastcone.py:272-292, cycles per pixel) divided by pw, so that
    FDK returns the attenuation per voxel (the reference's fbp() returns pw times it); the cosine
    weight is the cosine of the angle between each pixel's ray and the detector normal, from the
    geometry's view records (physical units; the reference's fbp() uses pixel units). D is not
    needed (the records carry the source) and is accepted for call compatibility. `out` may be
    `proj` itself (in place) when the dtypes match.
    """
    _need_device(proj, "proj")
    if tuple(proj.shape) != geom.proj_shape:
        raise ValueError(f"proj shape {tuple(proj.shape)} != geometry {geom.proj_shape}")
    if cosine and geom.kind != "flat":
        raise ValueError("the FDK cosine weight needs a flat-detector geometry")
    if out is None:
        out = torch.empty(proj.shape, dtype=out_dtype, device=proj.device)
    else:
        _need_device(out, "out")
        if tuple(out.shape) != tuple(proj.shape):
            raise ValueError("out must have the projection shape")
    views = geom.views_on(proj.device)
    L = _native.lib()
    rc = L.fc_fdk_filter_dev(proj.device.index, proj.data_ptr(), _dtype_code(proj, "proj"),
                             out.data_ptr(), _dtype_code(out, "out"), _proj3(geom),
                             views.data_ptr() if cosine else None, 1.0 / float(pw),
                             _stream(proj.device))
    _native.check(rc)
    return out


def fdk(proj, geom, D, pw, out=None, store_dtype=torch.float32):
    """Full-scan FDK: filter then FDK-weighted backprojection with norm = 1 / nviews."""
    q = fdk_filter(proj, geom, D, pw, out_dtype=store_dtype)
    norm = torch.full((geom.nviews,), 1.0 / geom.nviews, dtype=torch.float32, device=proj.device)
    return backproject(q, geom, _native.FC_BP_FDK, out=out, norm=norm)
