"""Build libgpu_tracking.so (gfx950) in-tree with hipcc.

Every HIP source is compiled separately (in parallel) and linked into
``tracking_optimizations_amd/libgpu_tracking.so``.  Flags:
  * ``-ffp-contract=off``: no implicit multiply-add contraction; the kernels
    write ``__builtin_fmaf`` where one fused operation is meant;
  * ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division and sqrt;
  * ``-fno-gpu-flush-denormals-to-zero``: keep binary32 subnormals;
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
    instructions are generated.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgpu_tracking.so")
SOURCES = ["gt_kernels.hip", "gt_synth.hip", "gt_adjoint.hip", "gt_link.hip", "gt_raw.hip", "gt_ingest.hip", "gt_api.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _compile(src: str, obj: str) -> None:
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")


def build(verbose: bool = True) -> str:
    """Compile every HIP source for gfx950 and link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s.replace(".hip", ".o")) for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        list(ex.map(lambda so: _compile(*so), zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libgpu_tracking.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


ASAN_DIR = os.path.join(HERE, "_build_asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libgpu_tracking_asan.so")


def build_asan(verbose: bool = False) -> str:
    """The library with its host code instrumented by AddressSanitizer (device
    code as usual, -fsanitize only after -Xarch_host), for CPU tests of the
    host paths (parameters, column names, movie-file decoding, argument and
    error handling) under clang's ASan runtime.  gt_kernels.hip, whose host
    code is kernel launchers only, is linked from the regular build's object
    (built first when missing)."""
    os.makedirs(ASAN_DIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in FLAGS if f != "-O3"] + [
        "-O1", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    jobs = []
    objs = []
    for src in SOURCES:
        if src == "gt_kernels.hip":
            obj = os.path.join(objdir, "gt_kernels.o")
            srcp = os.path.join(CSRC, src)
            if not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(srcp):
                jobs.append(([HIPCC, *FLAGS, "-c", srcp, "-o", obj], src))
        else:
            obj = os.path.join(ASAN_DIR, src.replace(".hip", ".o"))
            jobs.append(([HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj], src))
        objs.append(obj)

    def one(job):
        cmd, src = job
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc (asan) failed on {src}")

    with cf.ThreadPoolExecutor(max_workers=len(jobs) or 1) as ex:
        list(ex.map(one, jobs))
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host",
                        "-fsanitize=address", "-o", ASAN_LIB, *objs, "-lz"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of the asan library failed")
    if verbose:
        print(f"built {ASAN_LIB}")
    return ASAN_LIB


def asan_runtime() -> str:
    """clang's AddressSanitizer runtime matching hipcc (to preload)."""
    import glob
    cands = sorted(glob.glob(os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "llvm", "lib",
                                          "clang", "*", "lib", "linux",
                                          "libclang_rt.asan-x86_64.so")))
    cands += sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not cands:
        raise FileNotFoundError("libclang_rt.asan-x86_64.so not found")
    return cands[-1]


if __name__ == "__main__":
    build()
