"""Build steps of the package: the HIP library (gfx950) and the Fortran host.

Kept as plain command lines (no build system): hipcc compiles the single HIP source into
marfa_amd/libmarfa.so in place, so the built library travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmarfa.so")
HIP_SRC = [os.path.join(HERE, "csrc", "marfa_lbl.hip"), os.path.join(HERE, "csrc", "par_ingest.cpp")]
# every file the library is built from (sources, the kernel headers it includes, the C ABI)
HIP_DEPS = HIP_SRC + sorted(os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc"))
                            if f.endswith((".h", ".inc"))) + [os.path.join(ROOT, "include", "marfa.h")]

# packed FP32 instructions are switched off: the compiler emits ordinary VALU FP32 ops
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_hip(verbose: bool = False) -> str:
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", LIB + ".tmp"] + HIP_SRC
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    return LIB


TORCH_OPS_SRC = os.path.join(HERE, "csrc", "torch_ops.cpp")
TORCH_OPS_LIB = os.path.join(HERE, "libmarfa_torch.so")


def build_torch_ops(verbose: bool = False) -> str:
    """torch.ops.marfa.* (csrc/torch_ops.cpp): host code over the C ABI, compiled by hipcc
    against the installed PyTorch-ROCm and linked to libmarfa.so beside it."""
    import torch
    from torch.utils import cpp_extension as ce
    tlib = ce.library_paths()[0]
    cmd = [hipcc(), "-O2", "-std=c++17", "-fPIC", "-shared", "-DUSE_ROCM",
           f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}", "-Wno-unused-result"]
    cmd += ["-I" + d for d in ce.include_paths()]
    cmd += ["-o", TORCH_OPS_LIB + ".tmp", TORCH_OPS_SRC, "-L" + tlib, "-ltorch", "-ltorch_cpu", "-lc10",
            "-lc10_hip", "-ltorch_hip", "-L" + HERE, "-lmarfa", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + tlib]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc (torch operators) failed:\n" + r.stderr[-4000:])
    os.replace(TORCH_OPS_LIB + ".tmp", TORCH_OPS_LIB)
    return TORCH_OPS_LIB


def build_fortran_host(verbose: bool = False) -> str | None:
    """amdflang host programs marfa_amd/fortran/marfa_gpu (drop-in for app/main.f90) and
    marfa_amd/fortran/marfa_par (drop-in for app/processParFile.f90)."""
    src_dir = os.path.join(HERE, "fortran")
    srcs = [os.path.join(src_dir, f) for f in ("marfa_capi.f90", "marfa_main.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    out = os.path.join(src_dir, "marfa_gpu")
    mod = os.path.join(src_dir, "_mod")
    os.makedirs(mod, exist_ok=True)
    link = ["-L" + HERE, "-lmarfa", "-Wl,-rpath,$ORIGIN/.."]
    # marfa_gpu: drop-in for app/main.f90; marfa_par: drop-in for app/processParFile.f90
    for exe, files in ((out, srcs), (os.path.join(src_dir, "marfa_par"), [os.path.join(src_dir, "marfa_par.f90")])):
        cmd = [fc, "-O2", "-module-dir", mod, "-o", exe] + files + link
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed:\n" + r.stderr[-4000:])
    return out


def build_all(verbose: bool = False):
    build_hip(verbose)
    build_torch_ops(verbose)
    build_fortran_host(verbose)
