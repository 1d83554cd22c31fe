"""Build libpsx.so (the HIP/gfx950 kernels and the C ABI of include/psx.h) in-tree.

Each ``csrc/*.hip`` translation unit is compiled separately (in parallel) with
hipcc for gfx950 and the objects are linked into ``pyspectral_amd/libpsx.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libpsx.so")
ARCH = os.environ.get("PSX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: only FP64 is used, and their hand use is banned
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libpsx.so")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# per-source flags: the network kernels (and the z pass with the network fused in)
# keep matrix-core results in ordinary vector registers (no accumulation-register
# reads between the layers)
EXTRA = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"], "k_zlearn.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str, hipcc: str, build_dir: str, defines) -> str:
    obj = os.path.join(build_dir, os.path.basename(src)[:-4] + ".o")
    extra = EXTRA.get(os.path.basename(src), [])
    cmd = [hipcc, *FLAGS, *extra, *[f"-D{d}" for d in defines], "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = True, out: str = OUT, defines=()) -> str:
    """Compile every csrc/*.hip and link ``out``; ``defines`` are extra -D macros for
    experimental builds (compiled into a separate build folder)."""
    hipcc = _hipcc()
    build_dir = BUILD if not defines else BUILD + "_" + "_".join(d.replace("=", "") for d in defines)
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, hipcc, build_dir, defines), srcs))
    tmp = out + ".tmp"
    r = subprocess.run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} from {len(srcs)} sources")
    return out


HOSTTEST_OUT = os.path.join(HERE, "libpsx_hosttest.so")


def build_hosttest(verbose: bool = True) -> str:
    """Host-only library running the FFT butterflies and plans on the CPU
    (``csrc/host_fft_test.cpp``), used by the CPU test suite."""
    hipcc = _hipcc()
    src = os.path.join(CSRC, "host_fft_test.cpp")
    tmp = HOSTTEST_OUT + ".tmp"
    r = subprocess.run([hipcc, "-x", "hip", f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-fPIC", "-shared",
                        src, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host test library failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, HOSTTEST_OUT)
    if verbose:
        print(f"built {HOSTTEST_OUT}")
    return HOSTTEST_OUT


if __name__ == "__main__":
    build(jobs=int(sys.argv[1]) if len(sys.argv) > 1 else None)
    build_hosttest()
