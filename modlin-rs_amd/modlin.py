"""Python binding of libmodlin.so (the C ABI in include/modlin.h).

Mirrors the reference crate's public surface (src/lib.rs:9-14):

  Field(p)                       fp::Field::new            src/fp.rs:49-61
  Field.reduce / reduce_u128     Field::reduce[_u128]      src/fp.rs:63-79
  Context.rank(field, m)         fp::rank                  src/fp.rs:125
  Context.linear_complexity      fp::linear_complexity     src/fp.rs:260
  corank_prob, corank_tail_pvalue, lc_left_tail_pvalue, pretty_p_value, P_FLOOR
                                 stats::*                  src/stats.rs:16-137
  Prng                           prng::Prng                src/prng.rs

plus the batched entry points (rank_batch, lc_batch).  Errors the reference
turns into panics (p outside [2, 2^63), an empty BM sequence) raise ValueError
here; a missing library or GPU raises RuntimeError -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmodlin.so")

MODLIN_OK, MODLIN_EINVAL, MODLIN_ENOMEM, MODLIN_EHIP, MODLIN_ENODEV, MODLIN_EUNSUPPORTED = range(6)
METHOD_AUTO, METHOD_PLAIN, METHOD_MFMA = 0, 1, 2
ELEM_U64, ELEM_U32 = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
FLAG_PROFILE = 1
FLAG_NO_GRAPH = 2
FLAG_NO_DEFER = 4
FLAG_U64_STORE = 8
P_FLOOR = 1e-307  # stats.rs:16
LC_NOALLOC_MAX_N = 1 << 20  # MODLIN_LC_NOALLOC_MAX_N

PRNG_XOROSHIRO128PP, PRNG_MIXMAX17, PRNG_MIXMAX240, PRNG_MIXMAX256, PRNG_SPLITMIX64, PRNG_XOSHIRO256SS = range(6)

# Exported symbols, in the order of include/modlin.h.
EXPORTS = [
    "modlin_field_init", "modlin_reduce", "modlin_reduce_u128",
    "modlin_ctx_create", "modlin_ctx_destroy", "modlin_ctx_last_error", "modlin_ctx_last_stats",
    "modlin_rank", "modlin_linear_complexity", "modlin_rank_batch", "modlin_lc_batch",
    "modlin_linear_complexity_noalloc",
    "modlin_corank_prob", "modlin_corank_tail_pvalue", "modlin_lc_left_tail_pvalue", "modlin_pretty_p_value",
    "modlin_prng_fill", "modlin_prng_name", "modlin_prng_fill_device",
    "modlin_util_combine_rows", "modlin_util_lfsr",
    "modlin_dev_alloc", "modlin_dev_free", "modlin_memcpy_h2d", "modlin_memcpy_d2h", "modlin_synchronize",
    "modlin_ctx_set_tuning", "modlin_ctx_load_tuning", "modlin_tuning_select",
]


class _FieldT(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint64), ("mu_hi", ctypes.c_uint64), ("mu_lo", ctypes.c_uint64),
                ("safe_batch", ctypes.c_uint64)]


class _MatDesc(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("n", ctypes.c_uint64), ("elem", ctypes.c_uint32),
                ("mem", ctypes.c_uint32)]


class _Opts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_uint32), ("panel", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("coop_min", ctypes.c_uint32)]


TUNING_FIELDS = ("far_per_cu", "near_per_cu", "gather_wgs", "check_every", "panel", "bm_coop_wgs", "bm_threads",
                 "far_depth", "far_quarters")


class _Tuning(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in TUNING_FIELDS] + [("reserved", ctypes.c_uint32 * 3)]


class _Stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in
                ("macs", "update_ms", "update_launches", "limb_count", "iterations", "bm_bytes", "bm_ms",
                 "total_ms", "update_bytes", "graph", "bm_coop", "far_macs", "far_ms", "far_launches")]


_PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64)

_lib = None


def load() -> ctypes.CDLL:
    """Loads libmodlin.so (built in-tree); raises RuntimeError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB_PATH)
    u64, sz, vp = ctypes.c_uint64, ctypes.c_size_t, ctypes.c_void_p
    P64 = ctypes.POINTER(ctypes.c_uint64)
    lib.modlin_field_init.argtypes = [u64, ctypes.POINTER(_FieldT)]
    lib.modlin_reduce.argtypes = [ctypes.POINTER(_FieldT), u64]
    lib.modlin_reduce.restype = u64
    lib.modlin_reduce_u128.argtypes = [ctypes.POINTER(_FieldT), u64, u64]
    lib.modlin_reduce_u128.restype = u64
    lib.modlin_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.modlin_ctx_destroy.argtypes = [vp]
    lib.modlin_ctx_destroy.restype = None
    lib.modlin_ctx_last_error.argtypes = [vp]
    lib.modlin_ctx_last_error.restype = ctypes.c_char_p
    lib.modlin_ctx_last_stats.argtypes = [vp, ctypes.POINTER(_Stats)]
    lib.modlin_rank.argtypes = [vp, ctypes.POINTER(_FieldT), P64, u64, _PROGRESS, vp, P64]
    lib.modlin_linear_complexity.argtypes = [vp, ctypes.POINTER(_FieldT), P64, u64, _PROGRESS, vp, P64]
    lib.modlin_rank_batch.argtypes = [vp, ctypes.POINTER(_FieldT), ctypes.POINTER(_MatDesc), u64,
                                      ctypes.POINTER(_Opts), P64]
    lib.modlin_lc_batch.argtypes = [vp, ctypes.POINTER(_FieldT), ctypes.POINTER(_MatDesc), u64,
                                    ctypes.POINTER(_Opts), P64]
    lib.modlin_linear_complexity_noalloc.argtypes = [vp, ctypes.POINTER(_FieldT), vp, u64, P64]
    for name in ("modlin_corank_prob", "modlin_corank_tail_pvalue", "modlin_lc_left_tail_pvalue"):
        getattr(lib, name).argtypes = [u64, u64, u64]
        getattr(lib, name).restype = ctypes.c_double
    lib.modlin_pretty_p_value.argtypes = [ctypes.c_double, ctypes.c_char_p, sz]
    lib.modlin_prng_fill.argtypes = [ctypes.c_uint32, u64, u64, P64, u64]
    lib.modlin_prng_name.argtypes = [ctypes.c_uint32]
    lib.modlin_prng_name.restype = ctypes.c_char_p
    lib.modlin_prng_fill_device.argtypes = [vp, ctypes.c_uint32, u64, u64, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.POINTER(_FieldT), ctypes.c_uint32, vp, u64]
    lib.modlin_util_combine_rows.argtypes = [u64, P64, u64, u64, P64, u64, P64]
    lib.modlin_util_lfsr.argtypes = [u64, P64, u64, P64, u64, P64]
    lib.modlin_dev_alloc.argtypes = [vp, u64, ctypes.POINTER(vp)]
    lib.modlin_dev_free.argtypes = [vp, vp]
    lib.modlin_memcpy_h2d.argtypes = [vp, vp, vp, u64]
    lib.modlin_memcpy_d2h.argtypes = [vp, vp, vp, u64]
    lib.modlin_synchronize.argtypes = [vp]
    lib.modlin_ctx_set_tuning.argtypes = [vp, ctypes.POINTER(_Tuning)]
    lib.modlin_ctx_load_tuning.argtypes = [vp, ctypes.c_char_p]
    lib.modlin_tuning_select.argtypes = [ctypes.c_char_p, u64, u64, u64, ctypes.c_int, ctypes.POINTER(_Tuning)]
    _lib = lib
    return lib


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class Field:
    """A prime field F_p, p < 2^63 (src/fp.rs:34-61)."""

    def __init__(self, p: int):
        lib = load()
        self._f = _FieldT()
        if p < 2 or p >= 1 << 63 or lib.modlin_field_init(p, ctypes.byref(self._f)) != MODLIN_OK:
            raise ValueError("modulus must be at least 2 and below 2^63")
        self.p = p

    @property
    def mu(self) -> int:
        return (self._f.mu_hi << 64) | self._f.mu_lo

    @property
    def safe_batch(self) -> int:
        return self._f.safe_batch

    def reduce(self, x: int) -> int:
        return load().modlin_reduce(ctypes.byref(self._f), x)

    def reduce_u128(self, t: int) -> int:
        return load().modlin_reduce_u128(ctypes.byref(self._f), t >> 64, t & ((1 << 64) - 1))


class ModlinError(RuntimeError):
    pass


class Context:
    """A device context: stream, device arena, last-call statistics."""

    def __init__(self, device: int = 0):
        lib = load()
        h = ctypes.c_void_p()
        rc = lib.modlin_ctx_create(device, ctypes.byref(h))
        if rc != MODLIN_OK:
            raise RuntimeError(f"modlin_ctx_create failed with status {rc} (no gfx950 device?)")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.modlin_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc == MODLIN_OK:
            return
        msg = self._lib.modlin_ctx_last_error(self._h).decode(errors="replace")
        if rc == MODLIN_EINVAL:
            raise ValueError(f"{what}: {msg}")
        raise ModlinError(f"{what}: status {rc}: {msg}")

    # --- launch parameters (no effect on results) ------------------------
    def set_tuning(self, **fields):
        """Overrides the tuning table for later calls with the given nonzero
        fields (names in TUNING_FIELDS); no arguments removes the override."""
        if not fields:
            self._check(self._lib.modlin_ctx_set_tuning(self._h, None), "set_tuning")
            return
        t = _Tuning()
        for k, v in fields.items():
            if k not in TUNING_FIELDS:
                raise ValueError(f"unknown tuning field {k}")
            setattr(t, k, int(v))
        self._check(self._lib.modlin_ctx_set_tuning(self._h, ctypes.byref(t)), "set_tuning")

    def load_tuning(self, path: "str | None" = None):
        """Loads a tuning table (None: tuning.txt next to the library; "": none)."""
        self._check(self._lib.modlin_ctx_load_tuning(self._h, None if path is None else path.encode()),
                    "load_tuning")

    def stats(self) -> dict:
        s = _Stats()
        self._lib.modlin_ctx_last_stats(self._h, ctypes.byref(s))
        return {name: getattr(s, name) for name, _ in _Stats._fields_}

    # --- single-item entry points (src/fp.rs:125, 260) -------------------
    @staticmethod
    def _progress_fn(progress):
        """C callback for progress(done, total), or a null one."""
        if progress is None:
            return _PROGRESS()
        return _PROGRESS(lambda _user, done, total: progress(int(done), int(total)))

    def rank(self, field: Field, m: np.ndarray, n: int | None = None, progress=None) -> int:
        """Rank of the n x n matrix m; progress(done_columns, n) is called while it runs."""
        a = np.ascontiguousarray(m, dtype=np.uint64)
        if n is None:
            n = int(round(a.size ** 0.5))
        if a.size != n * n:
            raise ValueError("matrix must hold n*n entries")
        out = ctypes.c_uint64()
        ptr = _u64p(a) if a.size else None
        cb = self._progress_fn(progress)
        self._check(self._lib.modlin_rank(self._h, ctypes.byref(field._f), ptr, n, cb, None,
                                          ctypes.byref(out)), "rank")
        return out.value

    def linear_complexity(self, field: Field, s: np.ndarray, progress=None) -> int:
        """Linear complexity of s; progress(done_steps, len(s)) is called while it runs."""
        a = np.ascontiguousarray(s, dtype=np.uint64)
        out = ctypes.c_uint64()
        ptr = _u64p(a) if a.size else None
        cb = self._progress_fn(progress)
        self._check(self._lib.modlin_linear_complexity(self._h, ctypes.byref(field._f), ptr, a.size, cb,
                                                       None, ctypes.byref(out)), "linear_complexity")
        return out.value

    # --- batched entry points -------------------------------------------
    @staticmethod
    def _descs(items, square: bool):
        keep, descs = [], []
        for it in items:
            if isinstance(it, DeviceBuffer):
                descs.append(_MatDesc(it.ptr, it.n, it.elem, MEM_DEVICE))
                continue
            a = np.ascontiguousarray(it)
            if a.dtype == np.uint32:
                elem = ELEM_U32
            else:
                a = np.ascontiguousarray(a, dtype=np.uint64)
                elem = ELEM_U64
            n = int(round(a.size ** 0.5)) if square else a.size
            if square and n * n != a.size:
                raise ValueError("matrix must hold n*n entries")
            keep.append(a)
            descs.append(_MatDesc(a.ctypes.data if a.size else None, n, elem, MEM_HOST))
        arr = (_MatDesc * max(1, len(descs)))(*descs)
        return arr, keep

    def rank_batch(self, field: Field, mats: Sequence, method: int = METHOD_AUTO, panel: int = 0,
                   profile: bool = False, graph: bool = True, defer: bool = True, u32_store: bool = True) -> list[int]:
        """Ranks of a batch of square matrices.  defer=False updates the far columns
        once per panel instead of once per two panels; u32_store=False keeps a u64
        working copy for p < 2^32 (MFMA method; the same ranks either way)."""
        arr, keep = self._descs(mats, True)
        out = np.zeros(max(1, len(mats)), dtype=np.uint64)
        opts = _Opts(method, panel, (FLAG_PROFILE if profile else 0) | (0 if graph else FLAG_NO_GRAPH)
                     | (0 if defer else FLAG_NO_DEFER) | (0 if u32_store else FLAG_U64_STORE), 0)
        self._check(self._lib.modlin_rank_batch(self._h, ctypes.byref(field._f), arr, len(mats), ctypes.byref(opts),
                                                _u64p(out)), "rank_batch")
        return [int(x) for x in out[:len(mats)]]

    def lc_batch(self, field: Field, seqs: Sequence, profile: bool = False, coop_min: int = 0,
                 method: int = METHOD_AUTO) -> list[int]:
        """Linear complexities of a batch.  A batch of one sequence of length at
        least coop_min (0: the library default) runs cooperatively on the whole GPU;
        method=METHOD_PLAIN runs the one-step-per-pass kernel instead of blocks of
        64 steps (the result is the same)."""
        arr, keep = self._descs(seqs, False)
        out = np.zeros(max(1, len(seqs)), dtype=np.uint64)
        opts = _Opts(method, 0, FLAG_PROFILE if profile else 0, coop_min)
        self._check(self._lib.modlin_lc_batch(self._h, ctypes.byref(field._f), arr, len(seqs), ctypes.byref(opts),
                                              _u64p(out)), "lc_batch")
        return [int(x) for x in out[:len(seqs)]]

    def linear_complexity_noalloc(self, field: Field, seq: "DeviceBuffer") -> int:
        """modlin_linear_complexity_noalloc: the sequence in a u64 DeviceBuffer;
        the library allocates nothing during the call."""
        if seq.elem != ELEM_U64:
            raise ValueError("the allocation-free entry point takes u64 sequences")
        out = ctypes.c_uint64()
        self._check(self._lib.modlin_linear_complexity_noalloc(self._h, ctypes.byref(field._f), seq.ptr,
                                                               seq.nbytes // 8, ctypes.byref(out)),
                    "linear_complexity_noalloc")
        return out.value

    # --- device buffers ---------------------------------------------------
    def to_device(self, a: np.ndarray, square: bool = True) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        if a.dtype not in (np.uint32, np.uint64):
            a = a.astype(np.uint64)
        ptr = ctypes.c_void_p()
        self._check(self._lib.modlin_dev_alloc(self._h, a.nbytes, ctypes.byref(ptr)), "dev_alloc")
        self._check(self._lib.modlin_memcpy_h2d(self._h, ptr, a.ctypes.data, a.nbytes), "memcpy_h2d")
        n = int(round(a.size ** 0.5)) if square else a.size
        return DeviceBuffer(self, ptr.value, n, ELEM_U32 if a.dtype == np.uint32 else ELEM_U64, a.nbytes)

    def prng_fill_device(self, kind: int, seed: int, count: int, skip: int = 0, field: "Field | None" = None,
                         elem: int = ELEM_U64, keep: int = 0, drop: int = 0, square: bool = True) -> "DeviceBuffer":
        """Outputs skip .. skip+count-1 of a generator stream, reduced by `field`
        (raw next_u64() words when None), generated on the device
        (main.rs:94-105 fill loop; prng.rs:46-343)."""
        nbytes = count * (4 if elem == ELEM_U32 else 8)
        ptr = ctypes.c_void_p()
        self._check(self._lib.modlin_dev_alloc(self._h, nbytes, ctypes.byref(ptr)), "dev_alloc")
        fp = ctypes.byref(field._f) if field is not None else None
        rc = self._lib.modlin_prng_fill_device(self._h, kind, seed, skip, keep, drop, fp, elem, ptr, count)
        if rc != MODLIN_OK:
            self._lib.modlin_dev_free(self._h, ptr)
            self._check(rc, "prng_fill_device")
        n = int(round(count ** 0.5)) if square else count
        if square and n * n != count:
            n = count
        return DeviceBuffer(self, ptr.value, n, elem, nbytes)

    def synchronize(self):
        self._check(self._lib.modlin_synchronize(self._h), "synchronize")


class DeviceBuffer:
    """A device allocation holding one matrix or sequence (zero-copy input)."""

    def __init__(self, ctx: Context, ptr: int, n: int, elem: int, nbytes: int):
        self.ctx, self.ptr, self.n, self.elem, self.nbytes = ctx, ptr, n, elem, nbytes

    def to_host(self) -> np.ndarray:
        return self.read(0, self.nbytes // self.itemsize)

    @property
    def itemsize(self) -> int:
        return 4 if self.elem == ELEM_U32 else 8

    def read(self, offset: int, count: int) -> np.ndarray:
        """Elements offset .. offset+count-1, copied to the host."""
        if offset < 0 or count < 0 or (offset + count) * self.itemsize > self.nbytes:
            raise ValueError("read outside the buffer")
        a = np.empty(max(1, count), dtype=np.uint32 if self.elem == ELEM_U32 else np.uint64)
        self.ctx._check(self.ctx._lib.modlin_memcpy_d2h(self.ctx._h, a.ctypes.data, self.ptr + offset * self.itemsize,
                                                        count * self.itemsize), "memcpy_d2h")
        return a[:count]

    def write(self, offset: int, values: np.ndarray) -> None:
        """Stores values (this buffer's element type) at element offset."""
        a = np.ascontiguousarray(values, dtype=np.uint32 if self.elem == ELEM_U32 else np.uint64).ravel()
        if offset < 0 or (offset + a.size) * self.itemsize > self.nbytes:
            raise ValueError("write outside the buffer")
        self.ctx._check(self.ctx._lib.modlin_memcpy_h2d(self.ctx._h, self.ptr + offset * self.itemsize, a.ctypes.data,
                                                        a.nbytes), "memcpy_h2d")

    def view(self, offset: int, n: int) -> "DeviceBuffer":
        """A non-owning descriptor of the n x n matrix (or length-n sequence, for
        square=False callers) stored at element offset."""
        size = n * n
        if offset < 0 or (offset + size) * self.itemsize > self.nbytes:
            raise ValueError("view outside the buffer")
        v = DeviceBuffer(self.ctx, self.ptr + offset * self.itemsize, n, self.elem, size * self.itemsize)
        v.owner = False
        return v

    owner = True

    def free(self):
        if not self.owner:
            return
        if self.ptr:
            self.ctx._lib.modlin_dev_free(self.ctx._h, self.ptr)
            self.ptr = 0


# --- statistics (src/stats.rs) --------------------------------------------
def corank_prob(p: int, n: int, d: int) -> float:
    return load().modlin_corank_prob(p, n, d)


def corank_tail_pvalue(p: int, n: int, c: int) -> float:
    return load().modlin_corank_tail_pvalue(p, n, c)


def lc_left_tail_pvalue(p: int, n: int, ell: int) -> float:
    return load().modlin_lc_left_tail_pvalue(p, n, ell)


def pretty_p_value(x: float) -> str:
    buf = ctypes.create_string_buffer(64)
    if load().modlin_pretty_p_value(x, buf, 64) < 0:
        raise ValueError("buffer too small")
    return buf.value.decode()


# --- generators (src/prng.rs) ---------------------------------------------
def prng_fill(kind: int, seed: int, count: int, skip: int = 0) -> np.ndarray:
    out = np.zeros(max(1, count), dtype=np.uint64)
    rc = load().modlin_prng_fill(kind, seed, skip, _u64p(out), count)
    if rc != MODLIN_OK:
        raise ValueError("the zero seed is not allowed for MIXMAX" if rc == MODLIN_EINVAL else f"status {rc}")
    return out[:count]


def prng_name(kind: int) -> str:
    r = load().modlin_prng_name(kind)
    if r is None:
        raise ValueError("unknown generator")
    return r.decode()


# --- workload helpers (host) ----------------------------------------------
def combine_rows(p: int, rows: np.ndarray, coef: np.ndarray) -> np.ndarray:
    """out[i] = sum_k coef[i, k] * rows[k] mod p (exact)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    coef = np.ascontiguousarray(coef, dtype=np.uint64)
    nsrc, n = rows.shape
    nout = coef.shape[0]
    out = np.zeros((max(1, nout), max(1, n)), dtype=np.uint64)
    rc = load().modlin_util_combine_rows(p, _u64p(rows), nsrc, n, _u64p(coef), nout, _u64p(out))
    if rc != MODLIN_OK:
        raise ValueError("combine_rows: bad arguments")
    return out[:nout, :n]


def lfsr(p: int, coef: np.ndarray, init: np.ndarray, n: int) -> np.ndarray:
    """s[i] = sum_j coef[j] s[i-1-j] mod p for i >= L = len(coef), s[:L] = init."""
    coef = np.ascontiguousarray(coef, dtype=np.uint64)
    init = np.ascontiguousarray(init, dtype=np.uint64)
    out = np.zeros(max(1, n), dtype=np.uint64)
    rc = load().modlin_util_lfsr(p, _u64p(coef), coef.size, _u64p(init), n, _u64p(out))
    if rc != MODLIN_OK:
        raise ValueError("lfsr: bad arguments")
    return out[:n]


def tuning_select(path: str, p: int, max_n: int, count: int, bm: bool = False) -> dict:
    """The launch parameters a call (p, largest n, batch size) would take from the
    tuning table at `path` (zero: built-in default).  Needs no GPU."""
    t = _Tuning()
    rc = load().modlin_tuning_select(path.encode(), p, max_n, count, 1 if bm else 0, ctypes.byref(t))
    if rc != MODLIN_OK:
        raise ValueError(f"cannot use tuning table {path}")
    return {k: getattr(t, k) for k in TUNING_FIELDS}
