"""ctypes binding of libFastcone.so (built in-tree by fastcone_amd.build).

Loading fails loudly: there is no CPU fallback anywhere in the package.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libFastcone.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fastcone.h")

FC_OK, FC_ERR_ARGUMENT, FC_ERR_DEVICE, FC_ERR_HIP, FC_ERR_MEMORY = 0, 1, 2, 3, 4
FC_F32, FC_BF16, FC_F16 = 0, 1, 2
FC_BP_UNWEIGHTED, FC_BP_KATSEVICH, FC_BP_FDK = 0, 1, 2
FC_DET_FLAT, FC_DET_CYLINDER = 0, 1
FC_FILTER_MAX_COLS = 4096
FC_VIEW_DOUBLES = 32

_lib = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class FastconeError(RuntimeError):
    """A libFastcone call failed; `code` is the fc_status value."""

    def __init__(self, code, message):
        super().__init__(f"libFastcone error {code}: {message}")
        self.code = code


def _declare(lib):
    i, d, ll = ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    sig = {
        "fc_last_error": (i, []),
        "fc_last_error_message": (ctypes.c_char_p, []),
        "fc_clear_error": (None, []),
        "fc_version": (ctypes.c_char_p, []),
        "fc_device_count": (i, []),
        "fc_views_flat": (i, [i, _c_dbl_p, _c_dbl_p, d, d, d, _c_dbl_p, i, _c_int_p, _c_int_p,
                              _c_int_p, _c_dbl_p]),
        "fc_views_cylinder": (i, [i, _c_dbl_p, _c_dbl_p, d, d, d, d, _c_dbl_p, _c_dbl_p,
                                  _c_dbl_p]),
        "fc_project_dev": (i, [i, _vp, i, _c_int_p, _c_dbl_p, _vp, _c_int_p, _vp, d, _vp, _vp]),
        "fc_adjoint_dev": (i, [i, _vp, _c_int_p, _c_dbl_p, _vp, i, _c_int_p, _vp, d, _vp, _vp]),
        "fc_backproject_dev": (i, [i, _vp, _c_int_p, _c_dbl_p, _vp, i, _c_int_p, _vp, _vp, i,
                                   _vp]),
        "fc_sample_count_dev": (i, [i, _c_int_p, _c_dbl_p, _c_int_p, _vp, d, _vp, _vp]),
        "fc_convert_dev": (i, [i, _vp, i, _vp, i, ll, _vp]),
        "fc_fdk_filter_dev": (i, [i, _vp, i, _vp, i, _c_int_p, _vp, d, _vp]),
        "fc_fdk_filter_noalloc_dev": (i, [i, _vp, i, _vp, i, _c_int_p, _vp, d, _vp]),
        "fc_conv3d_fwd_dev": (i, [i, _vp, _vp, _vp, _vp, i, i, i, i, i, i, i, _vp]),
        "fc_conv3d_dgrad_dev": (i, [i, _vp, _vp, _vp, _vp, i, i, i, i, i, i, _vp]),
        "fc_conv3d_wgrad_dev": (i, [i, _vp, _vp, _vp, i, i, i, i, i, i, _vp]),
        "fc_mfma_layout_probe": (i, [i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "fc_katsevich_dev": (i, [i, _vp, i, i, i, i, d, d, d, d, d, d, d, d, d, d, i, _vp,
                                 _c_int_p, _c_dbl_p, _vp, _vp]),
        "fc_release_workspace": (i, [i]),
        "fc_set_tile_limits": (i, [ll, i]),
        "fc_profile_enable": (i, [i]),
        "fc_gather_bench": (i, [i, i, i, ctypes.POINTER(d)]),
        "fc_profile_reset": (i, []),
        "fc_profile_read": (i, [ctypes.c_char_p, ctypes.POINTER(ll), ctypes.POINTER(d)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    for name in ("GPU_CB_K_backproject", "GPU_CB_misalign_backproject", "GPU_CB_backproject",
                 "GPU_CB_misalign_project", "GPU_CB_project"):
        getattr(lib, name).restype = None


def lib():
    """Load libFastcone.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FastconeError(
                -1, f"{LIB_PATH} not found: build it with `python -m fastcone_amd.build`")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def check(rc=None):
    """Raise FastconeError if the last call on this thread failed."""
    L = lib()
    code = L.fc_last_error() if rc is None else rc
    if code != FC_OK:
        msg = L.fc_last_error_message().decode(errors="replace")
        raise FastconeError(code, msg)


def profile_enable(on=True):
    check(lib().fc_profile_enable(1 if on else 0))


def profile_reset():
    check(lib().fc_profile_reset())


def profile_read(kernel):
    """(launch count, total ms) of a kernel since the last reset."""
    c = ctypes.c_longlong(0)
    ms = ctypes.c_double(0.0)
    check(lib().fc_profile_read(kernel.encode(), ctypes.byref(c), ctypes.byref(ms)))
    return c.value, ms.value


def gather_bench(which, dtype=FC_F32, device=0):
    """Samples/s of the gather microbenchmark (see fc_gather_bench)."""
    r = ctypes.c_double(0.0)
    check(lib().fc_gather_bench(int(device), int(which), int(dtype), ctypes.byref(r)))
    return r.value


def header_functions():
    """Names of the functions declared in include/fastcone.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text,
                                 flags=re.M)) - {"if", "return"})


def int_array(vals):
    arr = (ctypes.c_int * len(vals))(*[int(v) for v in vals])
    return arr


def dbl_array(vals):
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])
