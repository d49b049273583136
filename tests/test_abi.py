"""The C-ABI library loads and exports every function include/modlin.h declares.

No GPU compute here: only symbol presence, host-side argument checks, and the
context constructor's refusal without a gfx950 device.
"""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT, modlin


def declared_functions():
    text = open(os.path.join(ROOT, "include", "modlin.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(modlin_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_match_binding():
    assert declared_functions() == sorted(modlin.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = modlin.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", modlin.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(modlin_\w+)", out))
    assert set(declared_functions()) <= exported


def test_header_compiles_as_c():
    src = '#include "modlin.h"\nint main(void){modlin_field_t f; return modlin_field_init(3,&f);}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_field_init_status_codes():
    lib = modlin.load()
    f = modlin._FieldT()
    assert lib.modlin_field_init(0, ctypes.byref(f)) == modlin.MODLIN_EINVAL
    assert lib.modlin_field_init(1 << 63, ctypes.byref(f)) == modlin.MODLIN_EINVAL
    assert lib.modlin_field_init((1 << 61) - 1, ctypes.byref(f)) == modlin.MODLIN_OK
    assert f.safe_batch == 64


def test_util_helpers_exact():
    p = (1 << 61) - 1
    import numpy as np
    rows = np.array([[p - 1, 2, 3], [5, p - 2, 7]], dtype=np.uint64)
    coef = np.array([[p - 1, 1], [2, 3]], dtype=np.uint64)
    out = modlin.combine_rows(p, rows, coef)
    want = [[(int(coef[i][0]) * int(rows[0][c]) + int(coef[i][1]) * int(rows[1][c])) % p for c in range(3)]
            for i in range(2)]
    assert out.tolist() == want
    s = modlin.lfsr(97, np.array([3, 5, 7]), np.array([1, 2, 3]), 10)
    ref = [1, 2, 3]
    for i in range(3, 10):
        ref.append((3 * ref[i - 1] + 5 * ref[i - 2] + 7 * ref[i - 3]) % 97)
    assert s.tolist() == ref
    # large primes: the accumulation must not overflow 128 bits
    for p in [(1 << 62) + 135, (1 << 63) - 25]:
        coef = np.array([p - 1 - i for i in range(40)], dtype=np.uint64)
        init = np.array([p - 2 - i for i in range(40)], dtype=np.uint64)
        s = modlin.lfsr(p, coef, init, 60)
        ref = [int(x) for x in init]
        for i in range(40, 60):
            ref.append(sum(int(coef[j]) * ref[i - 1 - j] for j in range(40)) % p)
        assert s.tolist() == ref


def test_context_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(RuntimeError):
        modlin.Context(0)


def test_reduce_invalid_field_sentinel():
    """modlin_reduce / modlin_reduce_u128 answer an invalid field with UINT64_MAX,
    a value no residue (< p < 2^63) can take; valid fields reduce as fp.rs:63-79."""
    lib = modlin.load()
    bad = modlin._FieldT()
    bad.p = 1
    assert lib.modlin_reduce(ctypes.byref(bad), 5) == (1 << 64) - 1
    assert lib.modlin_reduce_u128(ctypes.byref(bad), 1, 5) == (1 << 64) - 1
    assert lib.modlin_reduce(None, 5) == (1 << 64) - 1
    f = modlin.Field(97)
    assert f.reduce(1000) == 1000 % 97
    assert f.reduce_u128((1 << 127) + 3) == ((1 << 127) + 3) % 97
