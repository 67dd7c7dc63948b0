"""Build helpers for the gfx950 native code (runtime library and generated
group kernels).  Everything is compiled in-tree with hipcc; nothing is
JIT-compiled into a home-directory cache.
"""
import hashlib
import os
import shutil
import subprocess
import threading

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
TEMPLATE = os.path.join(PKG, "template", "hip")
INCLUDE = os.path.join(ROOT, "include")
LIB_PATH = os.path.join(PKG, "libbitgen_hip.so")
ARCH = os.environ.get("BITGEN_HIP_ARCH", "gfx950")

# Packed-FP32 instruction forms are disabled for every HIP translation unit.
# The atomic optimizer is off: it rewrites the tile-queue atomic so that its
# result is read right after it is sent, which defeats taking tiles ahead.
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
             "-std=c++17", "-Wno-unused-command-line-argument"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the hip backend)")


def build_runtime(force=True):
    """Compile libbitgen_hip.so (always recompiles when force is true)."""
    if not force and os.path.exists(LIB_PATH):
        return LIB_PATH
    src = os.path.join(CSRC, "bitgen_runtime.hip")
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-fPIC", "-shared", f"-I{INCLUDE}", f"-I{CSRC}", src, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def kernel_flags():
    return [*HIP_FLAGS, "--genco", f"-I{INCLUDE}", f"-I{TEMPLATE}", f"-I{CSRC}"]


def _dep_digest():
    h = hashlib.sha256()
    for path in (os.path.join(TEMPLATE, "bitgen_device.h"), os.path.join(INCLUDE, "bitgen_hip.h"),
                 os.path.join(CSRC, "transpose_core.h")):
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def source_key(source):
    """Cache key of a generated kernel: its text, the device headers, flags, arch."""
    h = hashlib.sha256()
    h.update(source.encode())
    h.update(_dep_digest().encode())
    h.update(" ".join(kernel_flags()).encode())
    return h.hexdigest()[:24]


def compile_kernel(src_path, out_path):
    """hipcc --genco one generated .hip file into a gfx950 code object.  The
    temporary name is unique per process and thread: two compiles of the same
    source (the cache is keyed by content) never write one file."""
    tmp = f"{out_path}.tmp{os.getpid()}_{threading.get_ident()}"
    cmd = [hipcc(), *kernel_flags(), src_path, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src_path}:\n{r.stderr[-4000:]}")
    os.replace(tmp, out_path)
    return out_path
