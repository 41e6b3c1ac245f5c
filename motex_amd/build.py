"""Builds the gfx950 library (motex_amd/lib/libmotex.so) and the motex command-line tool
(motex_amd/bin/motex) in-tree with hipcc.  ``python -m motex_amd.build`` or
``motex_amd.build.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libmotex.so")
BIN = os.path.join(HERE, "bin", "motex")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MOTEX_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
          "-Wall", "-Wno-unused-result"]
# Packed FP32 instructions are kept out of every device code object.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# Per-source device flags.  The Hamming kernels add each row's hit count to a per-row LDS word
# from the lanes that hit (a handful); the atomic optimizer would turn that uniform-address add
# into a scalar loop over the active lanes on every row.
HD_RING_SOURCES = [f"hd_ring_kernels_k{k}.hip" for k in (8, 10, 12, 14)]
SOURCE_FLAGS = {s: ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"] for s in ["hd_kernels.hip", *HD_RING_SOURCES]}
LIB_SOURCES = ["hd_kernels.hip", *HD_RING_SOURCES, "ed_kernels.hip", "ed_slice_kernels.hip", "ed_slice_kernels_k15.hip",
               "ed_slice_kernels_k23.hip", "struct_kernels.hip", "struct_range_kernels.hip", "struct_range_ring_short.hip",
               "struct_range_ring_long.hip", "runtime.hip", "runtime_ed.hip", "runtime_struct.hip", "runtime_serve.hip",
               "runtime_tune.hip", "runtime_noalloc.hip"]
CLI_SOURCES = ["cli/main.cpp", "cli/app.cpp", "cli/fasta.cpp", "cli/writer.cpp", "cli/writer_back.cpp", "cli/stats.cpp",
               "cli/switches.cpp"]
# MPFR (with GMP) for the command line's statistics modes (-b/-u/-U, -I), whose p-values the
# reference computes with 200-bit MPFR arithmetic: the system's runtime libraries (libmpfr6,
# libgmp10, which the compiler itself depends on) with the API headers of the conda
# installation.  Without them those modes are refused.
MPFR_INCLUDE = os.environ.get("MOTEX_MPFR_INCLUDE", "/opt/conda/include")
MPFR_LIBS = ["libmpfr.so.6", "libgmp.so.10"]
SYSTEM_LIB = "/usr/lib/x86_64-linux-gnu"


def mpfr_flags() -> tuple[list[str], list[str]]:
    """(compile flags, link flags) for MPFR, or empty lists when it is not installed."""
    if not (os.path.exists(os.path.join(MPFR_INCLUDE, "mpfr.h")) and os.path.exists(os.path.join(MPFR_INCLUDE, "gmp.h"))
            and all(os.path.exists(os.path.join(SYSTEM_LIB, l)) for l in MPFR_LIBS)):
        return [], []
    return ["-DMOTEX_HAVE_MPFR", f"-I{MPFR_INCLUDE}"], [f"-l:{l}" for l in MPFR_LIBS]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "_") + ".o")


def _compile_lib(src):
    out = _obj(src)
    _run([HIPCC, *COMMON, *DEVICE, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", out])
    return out


def _compile_cli(src):
    out = _obj(src)
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
          f"-I{CSRC}", *mpfr_flags()[0], "-c", os.path.join(CSRC, src), "-o", out])
    return out


def build(verbose: bool = False) -> tuple[str, str]:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    lib_srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cli_srcs = [s for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    workers = min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(workers) as ex:
        lib_objs = list(ex.map(_compile_lib, lib_srcs))
        cli_objs = list(ex.map(_compile_cli, cli_srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *lib_objs, "-lpthread"])
    os.replace(tmp, LIB)
    if cli_objs:
        _run(["g++", "-o", BIN, *cli_objs, f"-L{os.path.dirname(LIB)}", "-lmotex",
              "-Wl,-rpath,$ORIGIN/../lib", *mpfr_flags()[1], "-lpthread"])
    if verbose:
        print("built", LIB, BIN if cli_objs else "")
    return LIB, BIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
