"""Build recipe of the native parts, run in place (no install step):

* gturbo_amd/libgturbo_hip.so  -- the HIP kernels and the C ABI of
  include/gturbo.h, compiled for gfx950 with hipcc;
* gturbo_amd/fortran/gturbo_run -- the Fortran driver (amdflang), linked
  against the library through ISO_C_BINDING.

The kernels of each grid size are separate translation units compiled in
parallel.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgturbo_hip.so")
FDIR = os.path.join(HERE, "fortran")
FEXE = os.path.join(FDIR, "gturbo_run")

ARCH = "gfx950"
LOGNS = range(4, 14)
ENS_THREADS = (256, 512)  # workgroup sizes of the batched ensemble passes


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flang():
    for c in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
              shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def common_flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            f"-I{INCLUDE}", f"-I{CSRC}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_library(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()
    base = common_flags()
    src_k = os.path.join(CSRC, "gturbo_kernels.hip")
    units = []
    for t in ENS_THREADS:
        for sched in (0, 1):
            units.append((os.path.join(OBJ, f"kens{t}_{sched}.o"),
                          [src_k, f"-DGT_INST_ENS={t}", f"-DGT_ENS_SCHED={sched}"]))
    units.append((os.path.join(OBJ, "kcommon.o"), [src_k, "-DGT_INST_COMMON"]))
    units += [(os.path.join(OBJ, f"k{l}.o"), [src_k, f"-DGT_INST_LOGN={l}"]) for l in LOGNS]
    units.append((os.path.join(OBJ, "runtime.o"), [os.path.join(CSRC, "gturbo_runtime.cpp")]))
    units.append((os.path.join(OBJ, "ensemble.o"), [os.path.join(CSRC, "gturbo_ensemble.cpp")]))
    units.append((os.path.join(OBJ, "adjoint.o"), [os.path.join(CSRC, "gturbo_adjoint.cpp")]))
    jobs = jobs or min(16, os.cpu_count() or 4)

    def one(u):
        out, args = u
        _run([cc, *base, "-c", *args, "-o", out], verbose)
        return out

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, units))
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    """gturbo_run (the driver of src/main.f90) and gturbo_invert (initial-
    condition inversion through the adjoint), both on the library."""
    fc = flang()
    if fc is None:
        if verbose:
            print("no Fortran compiler: skipping the Fortran programs")
        return None
    common = ["gturbo_hip.f90", "gturbo_state.f90", "gturbo_sub.f90", "gturbo_io.f90"]
    programs = {"gturbo_run": "main.f90", "gturbo_invert": "invert.f90"}
    srcs = common + list(programs.values())
    if not all(os.path.exists(os.path.join(FDIR, s)) for s in srcs):
        return None
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = {}
    for src in srcs:  # in order: modules before their users
        obj = os.path.join(mod, src + ".o")
        _run([fc, "-O2", "-c", "-module-dir", mod, f"-I{mod}", os.path.join(FDIR, src), "-o",
              obj], verbose)
        objs[src] = obj
    for exe, main in programs.items():
        _run([fc, "-O2", *[objs[s] for s in common], objs[main], "-o",
              os.path.join(FDIR, exe), f"-L{HERE}", "-lgturbo_hip",
              "-Wl,-rpath,$ORIGIN/.."], verbose)
    return FEXE


def build(verbose=False):
    lib = build_library(verbose)
    exe = build_fortran(verbose)
    return lib, exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
