"""The AMReX adapters (include/fhdx_amrex.H, include/fhdx_amrex_stag.H) compile, call the C ABI of
fhdx.h with matching arguments, and define every reference entry point with the function type
the reference declares for it.

AMReX is not installed here, so tests/adapter/amrex_mock/ holds compile-only stand-ins for the
AMReX types and FHDeX namespaces the adapters read; tests/adapter/adapters_tu.cpp calls each
entry point with the reference's argument lists and takes its address as a pointer of the
reference's function type.  A planted parameter change must fail the same compile.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TU = os.path.join(ROOT, "tests", "adapter", "adapters_tu.cpp")


def _compile(*defines):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    cmd = [gxx, "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-variable", "-Werror", "-fsyntax-only",
           "-I", os.path.join(ROOT, "tests", "adapter", "amrex_mock"), "-I", os.path.join(ROOT, "include"),
           *[f"-D{d}" for d in defines], TU]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=120)


def test_adapters_compile_with_reference_signatures():
    r = _compile()
    assert r.returncode == 0, r.stderr


def test_planted_signature_mismatch_fails():
    r = _compile("PLANT_MISMATCH")
    assert r.returncode != 0
    assert "planted_p" in r.stderr
