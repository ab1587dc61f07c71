"""In-tree build of the product libraries (mocc_amd/lib) and the oracle.

``build_all()`` compiles:
  * libmocc_host.so  -- C++17 host front-end (g++),
  * libmocgpu.so     -- HIP for gfx950 (hipcc),
  * libmocc_h5.so    -- mocc's HDF5 output file (g++, the HDF5 C library),
  * mocc             -- the command-line program (g++),
  * oracle/_build/liboracle.so -- the FP64 checker (test infrastructure).
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mocc_amd")
LIB = os.path.join(PKG, "lib")
HOST_SRC = ["xml.cpp", "model.cpp", "problem.cpp", "host_api.cpp"]
GPU_SRC = ["plan.cpp", "sweep.hip", "cmfd.hip", "corr.hip", "mocgpu.hip", "batch.hip", "noalloc.hip"]
# No floating-point contraction in device code: a kernel is instantiated
# several ways (cells per lane of the coarse solve, with or without current
# tallies, single-problem or batched), and a batch picks the instantiation for
# all of its problems. Without contraction each multiply and add rounds on
# its own, so a problem's arithmetic does not depend on how the compiler fused
# a given instantiation; the hot loops write their fused multiply-adds
# explicitly.
GPU_EXTRA_FLAGS = {s: ["-ffp-contract=off"] for s in ("sweep.hip", "cmfd.hip", "corr.hip", "mocgpu.hip", "batch.hip")}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# No packed FP32 instructions (v_pk_*_f32) anywhere in device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_host():
    os.makedirs(LIB, exist_ok=True)
    src = [os.path.join(PKG, "csrc", "host", s) for s in HOST_SRC]
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-Wall",
          "-o", os.path.join(LIB, "libmocc_host.so.tmp")] + src)
    os.replace(os.path.join(LIB, "libmocc_host.so.tmp"), os.path.join(LIB, "libmocc_host.so"))


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(PKG, "csrc", "gpu", "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in GPU_SRC:
        src = os.path.join(PKG, "csrc", "gpu", s)
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        if s.endswith(".hip"):
            _run([HIPCC] + HIP_FLAGS + GPU_EXTRA_FLAGS.get(s, []) + ["-c", src, "-o", obj])
        else:
            _run([HIPCC, "-O3", "-std=c++17", "-fPIC", "-c", src, "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB, "libmocgpu.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)


# The HDF5 C library of this image (1.10); the output library is built only
# where its header is found. It is not put on the run path (that directory
# also holds an older libstdc++): callers load libhdf5 first by its full
# path (mocc_amd.output does).
HDF5_PREFIX = os.environ.get("MOCC_HDF5_PREFIX", "/opt/conda")


def build_h5():
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"libmocc_h5.so skipped: no hdf5.h under {inc}", file=sys.stderr)
        return
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmocc_h5.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I" + inc,
          os.path.join(PKG, "csrc", "h5", "output_h5.cpp"), "-o", out + ".tmp",
          "-L" + LIB, "-lmocc_host", "-L" + lib, "-lhdf5",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def _git_sha1() -> str:
    try:
        r = subprocess.run(["git", "-C", ROOT, "rev-parse", "HEAD"], capture_output=True, text=True)
        return r.stdout.strip() if r.returncode == 0 and r.stdout.strip() else "unknown"
    except OSError:
        return "unknown"


def build_cli():
    """The `mocc` command-line program (csrc/cli/mocc.cpp), next to the
    libraries it links (libmocgpu, libmocc_host; libmocc_h5 is loaded at run
    time after libhdf5)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "mocc")
    _run(["g++", "-std=c++17", "-O2", "-Wall", f'-DMOCC_GIT_SHA1="{_git_sha1()}"',
          os.path.join(PKG, "csrc", "cli", "mocc.cpp"), "-o", out + ".tmp",
          "-L" + LIB, "-lmocgpu", "-lmocc_host", "-ldl", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)


def build_all():
    build_host()
    build_gpu()
    build_h5()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    which = sys.argv[1:] or ["all"]
    for w in which:
        {"host": build_host, "gpu": build_gpu, "h5": build_h5, "cli": build_cli, "oracle": build_oracle,
         "all": build_all}[w]()
    print("built:", ", ".join(which))
