"""Build libfv2d_amd.so (HIP, gfx950) in-tree.

    python -m fv2d_amd.build [--jobs N]

The FP64 stage and the C-ABI layer are compiled with -ffp-contract=off (no a*b+c
contraction into FMA), which keeps the FP64 path bit-identical to the reference's
arithmetic. The FP32-form stage may contract. All HIP code is built with the packed-FP32
target feature disabled.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libfv2d_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# (source, extra flags)
UNITS = [
    ("stage_fp64.hip", ["-ffp-contract=off"]),
    ("stage_fp64_exact.hip", ["-ffp-contract=off"]),
    ("stage_fp32.hip", ["-ffp-contract=fast"]),
    ("fv2d_capi.hip", ["-ffp-contract=off"]),
    ("fv2d_batch.hip", ["-ffp-contract=off"]),
    ("fv2d_noalloc.hip", ["-ffp-contract=off"]),
]


def _compile(unit, defines=(), objdir=OBJ):
    src, extra = unit
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    dflags = [f"-D{d}" for d in defines]
    if any(d.startswith("FV2D_WAVES_PER_EU") for d in defines):
        extra = [x for x in extra if not x.startswith("-DFV2D_WAVES_PER_EU")]
    cmd = [HIPCC, *COMMON, *extra, *dflags, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that is harmless
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n" + "\n".join(noise))
    return obj, noise


DEFAULT_JOBS = max(1, min(len(UNITS), os.cpu_count() or 1))


def build(jobs: int = DEFAULT_JOBS, verbose: bool = False, defines=(), out: str = LIB) -> str:
    """Compile every unit and link `out`. `defines` (NAME=VALUE) select build variants."""
    objdir = OBJ if not defines else OBJ + "_" + "_".join(d.replace("=", "") for d in defines)
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(lambda u: _compile(u, defines, objdir), UNITS))
    objs = [o for o, _ in results]
    if verbose:
        for _, noise in results:
            for ln in noise:
                print(ln, file=sys.stderr)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, out)
    return out


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libfv2d_host.so")
HOST_EXE = os.path.join(HERE, "fv2d")
HOST_SOURCES = ["ini.cpp", "params.cpp", "init.cpp", "h5.cpp", "io.cpp", "host_capi.cpp"]
HDF5_INCLUDE = os.environ.get("FV2D_HDF5_INCLUDE", "/opt/conda/include")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-ffp-contract=off", f"-I{HDF5_INCLUDE}"]


def build_host(verbose: bool = False) -> tuple[str, str]:
    """The host side: libfv2d_host.so (.ini reader, parameters, initial conditions, HDF5
    snapshots; HDF5 is loaded at run time) and the fv2d executable (the time loop over the
    C ABI, linked against libfv2d_amd.so)."""
    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed: {' '.join(cmd)}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)

    tmp = HOST_LIB + ".tmp"
    run([CXX, *HOST_FLAGS, "-shared", "-o", tmp, *[os.path.join(HOST, s) for s in HOST_SOURCES], "-ldl"])
    os.replace(tmp, HOST_LIB)
    tmp = HOST_EXE + ".tmp"
    run([CXX, *HOST_FLAGS, "-o", tmp, os.path.join(HOST, "main.cpp"), os.path.join(HOST, "driver.cpp"),
         f"-L{HERE}", "-lfv2d_host", "-lfv2d_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_EXE)
    return HOST_LIB, HOST_EXE


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=DEFAULT_JOBS)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--define", action="append", default=[], help="NAME=VALUE for a build variant")
    ap.add_argument("--out", default=LIB)
    ap.add_argument("--host-only", action="store_true", help="build only libfv2d_host.so and fv2d")
    a = ap.parse_args(argv)
    if not a.host_only:
        print(build(a.jobs, a.verbose, tuple(a.define), a.out))
    if not a.define and a.out == LIB:
        print(*build_host(a.verbose))


if __name__ == "__main__":
    main()
