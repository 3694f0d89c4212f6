"""Build libpowspec_amd.so (gfx950) in-tree with hipcc.

The library is compiled from csrc/*.hip into lib/libpowspec_amd.so next to this file.
Every build compiles from scratch into a temporary directory and atomically replaces the
previous library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libpowspec_amd.so"
ARCH = "gfx950"
SOURCES = ["pk_capi.hip", "pk_scatter.hip", "pk_fft.hip", "pk_fused.hip", "pk_adjoint.hip", "pk_mock.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 ops stay off: hipcc emits them itself from float/float2 code otherwise
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    cc = hipcc()
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="pkbuild_")
    try:
        def compile_one(src):
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr)
            return obj

        jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            objs = list(ex.map(compile_one, SOURCES))
        out_tmp = os.path.join(tmp, LIBNAME)
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out_tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        final = lib_path()
        shutil.copy2(out_tmp, final + ".new")
        os.replace(final + ".new", final)
        return final
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose=True))
