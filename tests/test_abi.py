"""The C-ABI library loads and exports every entry point include/hzl_amd.h declares.

No compute calls: this runs without a GPU.
"""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "hzl_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hzl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_engine_surface():
    names = declared()
    assert "hzl_ensemble_create" in names and "hzl_ensemble_step" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol(pkg):
    path = pkg.engine.LIB_PATH
    if not os.path.exists(path):
        pytest.skip("engine library not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(pkg.EXPORTS) == declared()


def test_no_cpu_fallback_without_library(pkg, tmp_path):
    with pytest.raises(FileNotFoundError):
        pkg.engine.load_library.__wrapped__(str(tmp_path / "missing.so")) \
            if hasattr(pkg.engine.load_library, "__wrapped__") else \
            _fresh_load(pkg, str(tmp_path / "missing.so"))


def _fresh_load(pkg, path):
    saved = pkg.engine._lib
    pkg.engine._lib = None
    try:
        return pkg.engine.load_library(path)
    finally:
        pkg.engine._lib = saved


def test_header_compiles_as_c99_and_cxx():
    """The C-ABI header is plain C (cgo / ctypes / JNI callers include it from C)."""
    import shutil
    import subprocess
    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "include", "hzl_amd.h")
    for cmd in (["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                 "-x", "c", header],
                ["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", header]):
        if shutil.which(cmd[0]) is None:
            pytest.skip(f"{cmd[0]} not installed")
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
