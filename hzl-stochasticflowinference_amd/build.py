"""Build the engine's native library and the test infrastructure.

    python hzl-stochasticflowinference_amd/build.py        # everything

Products (all in-tree, git-ignored, shipped to the GPU host with the tree):
    hzl-stochasticflowinference_amd/libhzl_amd.so   C-ABI engine, HIP for gfx950
    hzl-stochasticflowinference_amd/libhzl_gpu.so   C++ drop-in (reference signatures)
    hzl-stochasticflowinference_amd/hzl_amd_cli     the reference CLI's ensemble / filter /
                                                    particle-sample commands on the GPU
Test infrastructure (never loaded by the product path):
    oracle/liboracle.so          FP64 C oracle
    oracle/_ref/libhzlref.so     unmodified reference CPU library (only where
                                 the reference sources are present)
    oracle/_ref/hzl_ref_cli      the reference's own CLI (same condition; writes the
                                 golden CLI outputs of tests/golden/cli)
    tests/emulator/libhzl_emu.so CPU lane emulator of the kernel bodies
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libhzl_amd.so")
EMU_SRC = os.path.join(ROOT, "tests", "emulator", "emu.cpp")
EMU_LIB = os.path.join(ROOT, "tests", "emulator", "libhzl_emu.so")
REFERENCE = os.environ.get("HZL_REFERENCE", "/root/reference")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 (v_pk_*_f32) is disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++20", "--offload-arch=gfx950", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]
SOURCES = [os.path.join(CSRC, f)
           for f in ("engine.hip", "variants.hip", "surrogate.hip", "enkf.hip", "pmcmc.hip",
                     "pf_multi.hip")]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_engine() -> str:
    """Compile every translation unit in parallel (one hipcc each), then link."""
    objs = [os.path.join(CSRC, os.path.basename(src) + ".o") for src in SOURCES]
    procs = []
    for src, obj in zip(SOURCES, objs):
        cmd = [HIPCC, *HIP_FLAGS, "-c", "-I", INCLUDE, "-I", CSRC, src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [src for src, p in zip(SOURCES, procs) if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, f"hipcc {' '.join(failed)}")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    return LIB


DROPIN_SRC = os.path.join(CSRC, "host", "hzl_gpu.cpp")
DROPIN_LIB = os.path.join(PKG, "libhzl_gpu.so")
DROPIN_CHECK_SRC = os.path.join(ROOT, "tests", "dropin", "dropin_check.cpp")
DROPIN_CHECK = os.path.join(ROOT, "tests", "dropin", "dropin_check")
CLI_SRC = os.path.join(CSRC, "host", "hzl_cli.cpp")
CLI = os.path.join(PKG, "hzl_amd_cli")


def build_dropin() -> str:
    """C++ drop-in layer (reference class signatures over the C-ABI) and the
    test driver that exercises it."""
    flags = ["-O2", "-std=c++20", "-Wall", "-I", INCLUDE, "-I", os.path.join(CSRC, "host")]
    tmp = DROPIN_LIB + ".tmp"
    _run(["g++", *flags, "-fPIC", "-shared", DROPIN_SRC, "-L", PKG, "-lhzl_amd",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, DROPIN_LIB)
    tmp = DROPIN_CHECK + ".tmp"
    _run(["g++", *flags, DROPIN_CHECK_SRC, "-L", PKG, "-lhzl_gpu", "-lhzl_amd",
          f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../hzl-stochasticflowinference_amd",
          "-o", tmp])
    os.replace(tmp, DROPIN_CHECK)
    tmp = CLI + ".tmp"
    _run(["g++", *flags, CLI_SRC, "-L", PKG, "-lhzl_gpu", "-lhzl_amd", "-Wl,-rpath,$ORIGIN",
          "-o", tmp])
    os.replace(tmp, CLI)
    return DROPIN_LIB


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    if os.path.isdir(os.path.join(REFERENCE, "src")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "ref-cli",
              f"REFERENCE={REFERENCE}"])


def build_emulator() -> str:
    tmp = EMU_LIB + ".tmp"
    _run([HIPCC, "-O2", "-std=c++20", "--offload-host-only", "-x", "hip", "-fPIC", "-shared",
          "-I", INCLUDE, "-I", CSRC, EMU_SRC, "-o", tmp])
    os.replace(tmp, EMU_LIB)
    return EMU_LIB


def build_all() -> None:
    # the CPU lane emulator (test infrastructure) compiles alongside the engine
    import threading
    errors = []

    def emulator():
        try:
            build_emulator()
        except Exception as e:  # reported after the engine build
            errors.append(e)

    t = threading.Thread(target=emulator)
    t.start()
    try:
        build_engine()
    finally:
        t.join()
    if errors:
        raise errors[0]
    build_dropin()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "engine": build_engine, "dropin": build_dropin,
         "oracle": build_oracle, "emulator": build_emulator}[w]()
