"""Exactness of the GPU arithmetic helpers, checked on the CPU.

tests/native/limb_check.cpp compiles the library's own header
(modlin-rs_amd/csrc/limb.hpp) for the host and compares the recombined residue
of int8 digit-panel products with the exact product mod p (128-bit integers),
on uniform panels and on the extreme panels that reach the class bounds.  It is
built once with g++ (portable carry code) and, when present, once with the ROCm
clang++ (the add-with-carry builtin the GPU code uses).  tests/native/fp_check.cpp
does the same for the field steps of fp_dev.hpp (the fused Mersenne step, the
fraction-free elimination step, the modular product) against 128-bit remainders.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "limb_check.cpp")
FP_SRC = os.path.join(ROOT, "tests", "native", "fp_check.cpp")
INC = os.path.join(ROOT, "modlin-rs_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"


def _compilers():
    out = []
    if shutil.which("g++"):
        out.append("g++")
    if os.path.exists(CLANG):
        out.append(CLANG)
    return out


@pytest.mark.parametrize("cxx", _compilers())
def test_limb_recombination_exact(cxx, tmp_path):
    exe = str(tmp_path / "limb_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", INC, "-o", exe, SRC], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("bad=0")


@pytest.mark.parametrize("cxx", _compilers())
def test_field_steps_exact(cxx, tmp_path):
    """lin2_m61, ff_step_t and mul_mod_t (fp_dev.hpp) against 128-bit remainders."""
    exe = str(tmp_path / "fp_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", INC, "-o", exe, FP_SRC], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("bad=0")
