"""Builds libmatpl_hip.so (the C-ABI library) in-tree for gfx950.

    python -m matpl_amd.build            # compile every source, link
    python -m matpl_amd.build --jobs 8

Each source under csrc/ is compiled with hipcc to an object in _build/, then
linked into matpl_amd/libmatpl_hip.so. Packed FP32 instructions are disabled
for every device pass (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmatpl_hip.so")
ARCH = os.environ.get("MATPL_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def cflags():
    # -amdgpu-mfma-vgpr-form: matrix-core accumulators in ordinary VGPRs (no AGPR round trips
    # between the MFMA layers and the activations around them)
    return ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-fvisibility=hidden",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form",
            "-Wno-unused-result", "-I", CSRC]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc(), *cflags(), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))   # not torch_ops/
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


TORCH_OPS_SRC = [os.path.join(CSRC, "torch_ops", "calcops_torch.cpp"),
                 os.path.join(CSRC, "torch_ops", "nep_ops_torch.cpp")]
TORCH_OPS_LIB = os.path.join(HERE, "libCalcOps_bind.so")


def build_torch_ops(verbose: bool = True) -> str:
    """libCalcOps_bind.so: the TORCH_LIBRARY registration of the ten CalcOps_cuda operators (the
    reference's library name) for torch.ops.load_library callers, TorchScript and LAMMPS hosts
    (host C++ against libtorch, linked to libmatpl_hip.so)."""
    import torch
    from torch.utils import cpp_extension as ce
    tdir = os.path.dirname(torch.__file__)
    inc = sum((["-I", p] for p in ce.include_paths()), [])
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
           "-DUSE_ROCM", f"-D_GLIBCXX_USE_CXX11_ABI={int(torch.compiled_with_cxx11_abi())}", *inc,
           "-I", "/opt/rocm/include", *TORCH_OPS_SRC, "-o", TORCH_OPS_LIB + ".tmp",
           "-L", os.path.join(tdir, "lib"), "-ltorch", "-ltorch_cpu", "-lc10", "-lc10_hip", "-ltorch_hip",
           "-L", HERE, "-lmatpl_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{os.path.join(tdir, 'lib')}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch-ops build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(TORCH_OPS_LIB + ".tmp", TORCH_OPS_LIB)
    if verbose:
        print(f"built {TORCH_OPS_LIB}", file=sys.stderr)
    return TORCH_OPS_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--torch-ops", action="store_true", help="also build libCalcOps_bind.so")
    a = ap.parse_args()
    build(a.jobs)
    if a.torch_ops:
        build_torch_ops()
