"""Build libFastcone.so in-tree for gfx950 (hipcc for HIP/C++ sources, gcc for wrapper.c).

    python -m fastcone_amd.build          # or fastcone_amd.build.build()

Every HIP translation unit is compiled with packed FP32 arithmetic disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops): the kernels are written in scalar FP32 and
must not be turned into v_pk_* FP32 instructions by the compiler.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libFastcone.so")
ARCH = os.environ.get("FASTCONE_ARCH", "gfx950")

HIP_SOURCES = ["fc_runtime.hip", "fc_project.hip", "fc_backproject.hip", "fc_util.hip",
               "fc_filter.hip",
               "fc_geometry.cpp", "fc_katsevich.hip", "fc_conv3d.hip"]
C_SOURCES = ["wrapper.c"]

HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
               "-munsafe-fp-atomics"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libFastcone.so)")


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose=False, jobs=None):
    """Compile every source and link fastcone_amd/libFastcone.so. Always recompiles."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    cmds, objs = [], []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([hipcc] + HIPCC_FLAGS + lang + ["-c", path, "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append(["gcc", "-O2", "-fPIC", "-std=c11", "-Wall", "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, len(cmds))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for err in ex.map(_compile, cmds):
            if verbose and err.strip():
                sys.stderr.write(err)
    tmp = LIB + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    _compile(link)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
