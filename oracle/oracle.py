"""ctypes binding of the CPU oracle (oracle/oracle.c) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py,
always as the checker.  The product (libmodlin.so) never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def build_native(out_dir: str) -> str:
    """The same library built for the host it runs on (-march=native), into
    out_dir: the CPU baseline's build, compiled where it is timed."""
    out = os.path.join(out_dir, "liboracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
                    "-o", out, os.path.join(_HERE, "oracle.c"), "-lm"], check=True)
    return out


def load(path: str | None = None) -> ctypes.CDLL:
    """The default library (built here for x86-64-v3), or the one at path."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    if path is None and not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(path or LIB_PATH)
    u64, sz, dbl = ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double
    P = ctypes.POINTER(ctypes.c_uint64)
    PD = ctypes.POINTER(ctypes.c_double)
    for name in ("or_field_mu_hi", "or_field_mu_lo", "or_field_safe_batch", "or_reduce"):
        getattr(lib, name).restype = u64
    lib.or_field_mu_hi.argtypes = lib.or_field_mu_lo.argtypes = lib.or_field_safe_batch.argtypes = [u64]
    lib.or_reduce.argtypes = [u64, u64]
    lib.or_reduce_u128.argtypes = [u64, u64, u64]
    lib.or_reduce_u128.restype = u64
    lib.or_mul.argtypes = [u64, u64, u64]
    lib.or_mul.restype = u64
    lib.or_inv.argtypes = [u64, u64]
    lib.or_inv.restype = u64
    lib.or_u128_mod.argtypes = [u64, u64, u64]
    lib.or_u128_mod.restype = u64
    lib.or_rank.argtypes = [u64, P, sz, P, PD]
    lib.or_rank.restype = sz
    lib.or_rank_ref.argtypes = [u64, P, sz]
    lib.or_rank_ref.restype = sz
    lib.or_rank_gf2.argtypes = [P, sz]
    lib.or_rank_gf2.restype = sz
    lib.or_lc.argtypes = [u64, P, sz, PD]
    lib.or_lc.restype = ctypes.c_longlong
    lib.or_lc_ref.argtypes = [u64, P, sz]
    lib.or_lc_ref.restype = ctypes.c_longlong
    for name in ("or_corank_prob", "or_corank_tail_pvalue", "or_lc_left_tail_pvalue"):
        getattr(lib, name).argtypes = [u64, sz, sz]
        getattr(lib, name).restype = dbl
    lib.or_pretty_p_value.argtypes = [dbl, ctypes.c_char_p, sz]
    lib.or_splitmix64_fill.argtypes = [u64, P, sz]
    lib.or_xoroshiro128pp_fill.argtypes = [u64, P, sz]
    lib.or_xoshiro256ss_fill.argtypes = [u64, P, sz]
    lib.or_mixmax_fill.argtypes = [ctypes.c_int, u64, u64, P, sz]
    lib.or_mixmax_fill.restype = ctypes.c_int
    lib.or_mix2.argtypes = lib.or_mix3.argtypes = [u64]
    lib.or_mix2.restype = lib.or_mix3.restype = u64
    lib.or_num_threads.restype = ctypes.c_int
    lib.or_set_num_threads.argtypes = [ctypes.c_int]
    if path is None:
        _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def field(p: int) -> dict:
    L = load()
    return {"p": p, "mu": (L.or_field_mu_hi(p) << 64) | L.or_field_mu_lo(p), "safe_batch": L.or_field_safe_batch(p)}


def reduce_u128(p: int, t: int) -> int:
    return load().or_reduce_u128(p, t >> 64, t & ((1 << 64) - 1))


def reduce(p: int, x: int) -> int:
    return load().or_reduce(p, x)


def mul(p: int, a: int, b: int) -> int:
    return load().or_mul(p, a, b)


def inv(p: int, a: int) -> int:
    return load().or_inv(p, a)


def rank(p: int, m: np.ndarray, n: int, with_trace: bool = False, lib: ctypes.CDLL | None = None):
    a = np.array(m, dtype=np.uint64, copy=True).reshape(-1)
    pc = np.zeros(max(1, n), dtype=np.uint64)
    w = ctypes.c_double()
    r = (lib or load()).or_rank(p, _p(a), n, _p(pc), ctypes.byref(w))
    if with_trace:
        return r, [int(x) for x in pc[:r]], w.value
    return r


def rank_ref(p: int, m: np.ndarray, n: int) -> int:
    a = np.array(m, dtype=np.uint64, copy=True).reshape(-1)
    return load().or_rank_ref(p, _p(a), n)


def rank_gf2(m: np.ndarray, n: int) -> int:
    a = np.ascontiguousarray(m, dtype=np.uint64).reshape(-1)
    return load().or_rank_gf2(_p(a), n)


def lc(p: int, s: np.ndarray, with_stats: bool = False, lib: ctypes.CDLL | None = None):
    a = np.ascontiguousarray(s, dtype=np.uint64)
    st = (ctypes.c_double * 3)()
    r = (lib or load()).or_lc(p, _p(a), a.size, st)
    if r < 0:
        raise ValueError("empty sequence")
    if with_stats:
        return r, {"bytes": st[0], "mulmods": st[1], "changes": st[2]}
    return r


def lc_ref(p: int, s: np.ndarray) -> int:
    a = np.ascontiguousarray(s, dtype=np.uint64)
    return load().or_lc_ref(p, _p(a), a.size)


def corank_prob(p, n, d):
    return load().or_corank_prob(p, n, d)


def corank_tail_pvalue(p, n, c):
    return load().or_corank_tail_pvalue(p, n, c)


def lc_left_tail_pvalue(p, n, ell):
    return load().or_lc_left_tail_pvalue(p, n, ell)


def pretty_p_value(x: float) -> str:
    buf = ctypes.create_string_buffer(96)
    load().or_pretty_p_value(x, buf, 96)
    return buf.value.decode()


def splitmix64(seed: int, count: int) -> np.ndarray:
    o = np.zeros(max(1, count), dtype=np.uint64)
    load().or_splitmix64_fill(seed, _p(o), count)
    return o[:count]


def xoroshiro128pp(seed: int, count: int) -> np.ndarray:
    o = np.zeros(max(1, count), dtype=np.uint64)
    load().or_xoroshiro128pp_fill(seed, _p(o), count)
    return o[:count]


def xoshiro256ss(seed: int, count: int) -> np.ndarray:
    o = np.zeros(max(1, count), dtype=np.uint64)
    load().or_xoshiro256ss_fill(seed, _p(o), count)
    return o[:count]


def mixmax(variant: int, seed: int, count: int, skip: int = 0) -> np.ndarray:
    o = np.zeros(max(1, count), dtype=np.uint64)
    if load().or_mixmax_fill(variant, seed, skip, _p(o), count) != 0:
        raise ValueError("the zero seed is not allowed for MIXMAX")
    return o[:count]


def mix2(x: int) -> int:
    return load().or_mix2(x)


def mix3(x: int) -> int:
    return load().or_mix3(x)


def num_threads() -> int:
    return load().or_num_threads()


def set_num_threads(t: int) -> None:
    load().or_set_num_threads(t)
