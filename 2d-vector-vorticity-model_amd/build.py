"""Build libvvm_amd.so in-tree: every csrc/*.hip (device + C ABI) and
csrc/*.cpp (host) translation unit through hipcc for gfx950, then one
shared-library link.  Run as a script or through __graft_entry__.build()."""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VVM_OFFLOAD_ARCH", "gfx950")
OUT = os.path.join(PKG, "libvvm_amd.so")
OBJ = os.path.join(PKG, "_build")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{ROOT}/include", f"-I{PKG}/csrc", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-value",
          "-Wno-unused-result"]
# Device code: gfx950 only; packed-FP32 instructions off; the FP64 model
# kernels keep the reference's rounding (no FMA contraction), set per file
# with #pragma clang fp contract where needed.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip"] + DEVICE + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):  # plain host code
        cmd = [os.environ.get("CXX", "g++")] + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = (r.stderr or "").strip()
    if warn:
        lines = [l for l in warn.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")) + glob.glob(os.path.join(PKG, "csrc", "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_host()
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources, {HOST_LIB} and {HOST_MAIN}")
    return OUT


HOST_LIB = os.path.join(PKG, "libvvm.so")
HOST_MAIN = os.path.join(PKG, "vvm_main")


def build_host() -> None:
    """The C++ state/config API (host/vvm.cpp -> libvvm.so) and the model
    driver (host/vvm_main.cpp -> vvm_main), both plain host code over
    libvvm_amd.so."""
    cxx = os.environ.get("CXX", "g++")
    base = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT}/include"]
    link = [f"-L{PKG}", "-lvvm_amd", "-Wl,-rpath,$ORIGIN"]
    steps = ((base + ["-fPIC", "-shared", os.path.join(PKG, "host", "vvm.cpp"), "-o", HOST_LIB + ".tmp"] + link,
              HOST_LIB),
             (base + [os.path.join(PKG, "host", "vvm_main.cpp"), "-o", HOST_MAIN + ".tmp", "-lvvm"] + link,
              HOST_MAIN))
    for cmd, target in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if r.stderr.strip():
            sys.stderr.write(r.stderr)
        os.replace(target + ".tmp", target)


if __name__ == "__main__":
    build(verbose=True)
