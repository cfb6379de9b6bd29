"""Builds mpart_amd/libmpart_amd.so from mpart_amd/csrc/*.hip for gfx950.

Usage: ``python mpart_amd/build.py [--incremental]`` (also called by
``__graft_entry__.build()``, which compiles everything).
Objects go to mpart_amd/_build/; the translation units compile in parallel.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpart_amd.so")

ARCH = os.environ.get("MPART_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC",
    # Keep the compiler from forming packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

SOURCES = ["mono_host.hip", "codegen.hip", "k_eval.hip", "k_coef.hip", "k_input.hip", "k_inverse.hip", "k_inverse_batch.hip",
           "k_fused.hip", "k_batch.hip", "k_group.hip",
           "k_group_fused.hip", "k_split.hip", "k_split2.hip", "k_split_eval.hip"]
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")


def embed_device_header():
    """Writes csrc/_mono_device_src.inc: mono_device.hpp as a C++ raw string
    literal, compiled into the library for the runtime (hipRTC) specialization."""
    with open(os.path.join(CSRC, "mono_device.hpp")) as f:
        text = f.read()
    if ")MPDSRC\"" in text:
        raise RuntimeError("mono_device.hpp contains the raw-string delimiter")
    out = os.path.join(CSRC, "_mono_device_src.inc")
    body = 'R"MPDSRC(' + text + ')MPDSRC"\n'
    old = open(out).read() if os.path.exists(out) else None
    if old != body:
        with open(out, "w") as f:
            f.write(body)


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def _stale(src):
    """An object is stale when it is missing or older than its source or any
    header of csrc/ (headers are shared by every translation unit)."""
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC)
                                        if f.endswith((".hpp", ".inc"))]
    deps.append(os.path.join(HERE, "..", "include", "mpart_amd.h"))
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build(verbose=True, jobs=None, incremental=False):
    os.makedirs(BUILD, exist_ok=True)
    embed_device_header()
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    todo = [src for src in SOURCES if not incremental or _stale(src)]
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(BUILD, os.path.splitext(src)[0] + ".o") for src in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, f"-L{ROCM_LIB}", "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    # --incremental recompiles only stale objects (for development; the
    # default, and __graft_entry__.build(), compiles everything)
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
