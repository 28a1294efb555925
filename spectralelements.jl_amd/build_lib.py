"""Builds libsemhip.so (gfx950) from csrc/ with hipcc, in-tree.

Usage: python build_lib.py [--jobs N]. Objects go to csrc/_build/, the library
to lib/libsemhip.so.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsemhip.so")

# (source, extra flags, object name); the fused apply is instantiated once per
# P in its own unit so the instantiations compile in parallel
UNITS = ([("sem_host.cpp", [], "sem_host"), ("k_ax.hip", [], "k_ax"),
          ("k_misc.hip", [], "k_misc"), ("k_pcg.hip", [], "k_pcg"),
          ("k_mg.hip", [], "k_mg"), ("k_geom.hip", [], "k_geom"),
          ("k_generic.hip", [], "k_generic"), ("sem_cd.cpp", [], "sem_cd")]
         + [("k_ax_inst.hip", [f"-DSEM_P={p}", f"-DSEM_PART={q}"], f"k_ax_p{p}_{q}")
            for p in range(16, 1, -1) for q in (range(7) if p == 16 else range(2))])

# Packed FP32 instructions are disabled: some of their forms have produced
# wrong results on these parts; scalar FP32 code is used instead.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def gen_tables(verbose: bool = False) -> None:
    """Compiles and runs gen_dtables.cpp (host g++), writing the GLL
    derivative tables the fused kernels hold as compile-time constants."""
    gen_dir = os.path.join(CSRC, "_gen")
    os.makedirs(gen_dir, exist_ok=True)
    exe = os.path.join(BUILD, "gen_dtables")
    cxx = shutil.which("g++") or shutil.which("c++")
    r = subprocess.run([cxx, "-O2", "-std=c++17", os.path.join(CSRC, "gen_dtables.cpp"), "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gen_dtables build failed:\n{r.stderr}")
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    path = os.path.join(gen_dir, "gll_dtables.h")
    old = open(path).read() if os.path.exists(path) else None
    if old != out:
        with open(path, "w") as f:
            f.write(out)
    if verbose:
        print(f"tables -> {path}")


def _newest_input() -> float:
    t = 0.0
    for root, _, files in os.walk(CSRC):
        if os.path.basename(root) == "_build":
            continue
        for f in files:
            if f.endswith((".hip", ".cpp", ".h")):
                t = max(t, os.path.getmtime(os.path.join(root, f)))
    return max(t, os.path.getmtime(os.path.join(os.path.dirname(HERE), "include", "sem_abi.h")))


def build(jobs: int = 8, verbose: bool = False, force: bool = True, defines=(),
          lib: str = None, objdir: str = None) -> str:
    """Compiles every unit (force=True) or only units older than any source.
    defines (e.g. ["-DSEM_AX_THREADS=128"]) are added to every hipcc line;
    lib / objdir redirect the library and the objects (tuning builds)."""
    out_lib = lib or LIB
    bdir = objdir or BUILD
    os.makedirs(bdir, exist_ok=True)
    os.makedirs(os.path.dirname(out_lib), exist_ok=True)
    gen_tables(verbose)
    cc = hipcc()
    newest = _newest_input()
    objs = []
    t0 = time.time()
    pending = []
    for s, extra, name in UNITS:
        o = os.path.join(bdir, name + ".o")
        objs.append(o)
        if force or not os.path.exists(o) or os.path.getmtime(o) < newest:
            pending.append((s, extra, name))
    running = []
    while pending or running:
        while pending and len(running) < jobs:
            s, extra, name = pending.pop(0)
            o = os.path.join(bdir, name + ".o")
            cmd = [cc, *FLAGS, *defines, *extra, "-x", "hip", "-c", os.path.join(CSRC, s), "-o", o]
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        still = []
        for s, p in running:
            if p.poll() is None:
                still.append((s, p))
                continue
            out = p.stdout.read().decode(errors="replace")
            if p.returncode != 0:
                for _, q in still:
                    q.kill()
                raise RuntimeError(f"hipcc failed on {s}:\n{out}")
            if verbose and out.strip():
                print(out)
        running = still
        time.sleep(0.05)
    tmp = out_lib + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out_lib)
    if verbose:
        print(f"built {out_lib} in {time.time() - t0:.1f}s")
    return out_lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only units older than the newest source")
    a = ap.parse_args()
    try:
        build(a.jobs, a.verbose, force=not a.incremental)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
