"""Build recipe for the engine library and the command-line driver.

Compiles, for gfx950 only:
  periodsearch_amd/libpsamd.so        -- csrc/ps_engine.hip + csrc/ps_mlp.hip + csrc/ps_shape.hip
                                         + csrc/ps_host.cpp + csrc/ps_learned.cpp
                                         (C ABI of include/periodsearch.h and include/ps_proposal.h)
  periodsearch_amd/period_search_amd  -- csrc/ps_main.cpp linked against libpsamd.so

Run ``python periodsearch_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libpsamd.so")
EXE = os.path.join(HERE, "period_search_amd")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions stay disabled in every device pass.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd: list[str]) -> None:
    t0 = time.time()
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")
    dt = time.time() - t0
    print(f"[build] {os.path.basename(cmd[-1]) if cmd else ''} ok in {dt:.1f}s", flush=True)


def build(verbose: bool = False, defines=(), lib_out: str | None = None, driver: bool = True) -> str:
    """Compile the library (and the driver).  ``defines`` are extra -D flags
    for kernel variants; ``lib_out`` writes the library elsewhere."""
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    tag = "_".join(d.replace("=", "") for d in defines)
    eng_o = os.path.join(HERE, "_obj", f"ps_engine{tag}.o")
    host_o = os.path.join(HERE, "_obj", "ps_host.o")
    mlp_o = os.path.join(HERE, "_obj", "ps_mlp.o")
    shape_o = os.path.join(HERE, "_obj", "ps_shape.o")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, *DEVICE_FLAGS,
          *[f"-D{d}" for d in defines], "-c", "-o", eng_o, os.path.join(CSRC, "ps_engine.hip")])
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, *DEVICE_FLAGS,
          "-c", "-o", mlp_o, os.path.join(CSRC, "ps_mlp.hip")])
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, *DEVICE_FLAGS,
          "-c", "-o", shape_o, os.path.join(CSRC, "ps_shape.hip")])
    # no sin/cos fusion by the compiler (see csrc/ps_host.cpp)
    _run(["g++", "-O3", "-fPIC", "-std=c++17", "-fno-builtin-sin", "-fno-builtin-cos", "-fno-builtin-sincos",
          "-I", INCLUDE, "-c", "-o", host_o,
          os.path.join(CSRC, "ps_host.cpp")])
    learned_o = os.path.join(HERE, "_obj", "ps_learned.o")
    _run(["g++", "-O2", "-fPIC", "-std=c++17", "-pthread", "-I", INCLUDE, "-c", "-o", learned_o,
          os.path.join(CSRC, "ps_learned.cpp")])
    objs = [eng_o, mlp_o, shape_o, host_o, learned_o]
    lib = lib_out or LIB
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, lib)
    if not driver:
        return lib
    _run(["g++", "-O2", "-std=c++17", "-I", INCLUDE, "-o", EXE, os.path.join(CSRC, "ps_main.cpp"),
          "-L", HERE, "-lpsamd", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    return lib


if __name__ == "__main__":
    build()
