"""Build the native library in-tree: ctpwa_amd/libctpwa_amd.so (gfx950), and the PyTorch
extension over its C ABI, ctpwa_amd/_ctpwa_torch.so (the likelihood as a C++ autograd
function; host code only, compiled with g++ against torch's headers).

    python -m ctpwa_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libctpwa_amd.so")
TORCH_EXT = os.path.join(HERE, "_ctpwa_torch.so")
SOURCES = ["init.hip", "prep.hip", "hessian.hip", "capi.hip", "codegen.cpp", "codegen_hess.cpp"]
ARCH = os.environ.get("CTPWA_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    objs = []
    flags = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
    # keep the compiler from forming packed fp32 instructions in device code
    dev_flags = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(HERE, "_build", s + ".o")
        cmd = [hipcc()] + flags
        if s.endswith(".hip"):
            cmd += dev_flags + ["-x", "hip"]
        cmd += ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_torch_ext(verbose)
    return LIB


def build_torch_ext(verbose: bool = False) -> str:
    """_ctpwa_torch.so: links libctpwa_amd.so from its own directory and torch's libraries."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    inc = ce.include_paths() + [sysconfig.get_paths()["include"]]
    libdirs = ce.library_paths()
    tmp = TORCH_EXT + ".tmp"
    cmd = (["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
            "-D_GLIBCXX_USE_CXX11_ABI=%d" % int(torch._C._GLIBCXX_USE_CXX11_ABI),
            "-DTORCH_EXTENSION_NAME=_ctpwa_torch", "-DTORCH_API_INCLUDE_EXTENSION_H",
            os.path.join(CSRC, "torch_ops.cpp"), "-o", tmp]
           + ["-I" + i for i in inc] + ["-L" + HERE, "-lctpwa_amd", "-Wl,-rpath,$ORIGIN"]
           + ["-L" + d for d in libdirs] + ["-Wl,-rpath," + d for d in libdirs]
           + ["-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python"])
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, TORCH_EXT)
    return TORCH_EXT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
