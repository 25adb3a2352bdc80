"""ctypes binding of libjlc_hip.so (the C ABI declared in include/jlc_hip.h).

This is the Python analogue of the Java FFM downcall layer a JLC maintainer
adds beside GemmDispatch: plain pointers and sizes, an int status per call,
and the thread-local error message turned into an exception. There is no
fallback: if the shared library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libjlc_hip.so"
# The autotuner's table for MI355X; loaded with the library when present.
TUNE_TABLE = Path(__file__).resolve().parent / "tuning" / "mi355x.tune"

JLC_OK = 0
JLC_ERR_INVALID_ARG = 1
JLC_ERR_NO_DEVICE = 2
JLC_ERR_DISABLED = 3
JLC_ERR_HIP = 4
JLC_ERR_ALLOC = 5
JLC_ERR_UNSUPPORTED = 6
JLC_ERR_ACCURACY = 7

ROW = 0
COL = 1

MODE_FP64 = 0
MODE_SPLIT_INT8 = 1
MODE_SPLIT_BF16 = 2
MODE_SPLIT_CRT = 3

KERNEL_NAMES = {0: "none", 1: "scale", 2: "mfma_f64", 3: "gemv", 4: "split_int8", 5: "split_bf16", 6: "grouped",
                7: "qr", 8: "split_crt", 9: "wy"}


def mode_slices(mode: int, slices: int) -> int:
    return mode | (slices << 8)


class JlcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"jlc status {status}: {message}")
        self.status = status
        self.message = message


class NativeLibraryMissing(ImportError):
    pass


class Stats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("kernel", C.c_int32),
        ("tile_m", C.c_int32),
        ("tile_n", C.c_int32),
        ("tile_k", C.c_int32),
        ("slices", C.c_int32),
        ("launches", C.c_int32),
        ("bits_a", C.c_int32),
        ("bits_b", C.c_int32),
        ("main_ms", C.c_double),
    ]

    @property
    def kernel_name(self) -> str:
        return KERNEL_NAMES.get(self.kernel, str(self.kernel))


_i64 = C.c_int64
_d = C.c_double
_p = C.c_void_p
_i = C.c_int

class GemmDesc(C.Structure):
    """jlc_gemm_desc: one problem of a grouped launch (device pointers)."""
    _fields_ = [
        ("layout_a", C.c_int32), ("layout_b", C.c_int32),
        ("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32), ("flags", C.c_int32),
        ("alpha", C.c_double), ("beta", C.c_double),
        ("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p),
        ("lda", C.c_int64), ("ldb", C.c_int64), ("ldc", C.c_int64),
    ]


class QrDesc(C.Structure):
    """jlc_qr_desc: one matrix of a batched QR (device pointers, column-major)."""
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("lda", C.c_int64), ("a", C.c_void_p), ("tau", C.c_void_p)]


_SIGNATURES = {
    "jlc_hip_probe": (_i, [C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "jlc_hip_init": (_i, [_i]),
    "jlc_hip_shutdown": (_i, []),
    "jlc_last_error": (C.c_char_p, []),
    "jlc_version": (C.c_char_p, []),
    "jlc_dgemm": (_i, [_i, _i, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _i, C.POINTER(Stats)]),
    "jlc_dgemm_dev": (_i, [_i, _i, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _i, _p,
                           C.POINTER(Stats)]),
    "jlc_dgemm_dev_noalloc": (_i, [_i, _i, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p]),
    "jlc_wy_apply_dev": (_i, [_i, _i, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _i, _p, C.POINTER(Stats)]),
    "jlc_wy_apply": (_i, [_i, _i, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _i, C.POINTER(Stats)]),
    "jlc_dgemv": (_i, [_i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, C.POINTER(Stats)]),
    "jlc_dgemv_dev": (_i, [_i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p, C.POINTER(Stats)]),
    "jlc_dev_alloc": (_i, [C.POINTER(_p), C.c_size_t]),
    "jlc_dev_free": (_i, [_p]),
    "jlc_memcpy_h2d": (_i, [_p, _p, C.c_size_t]),
    "jlc_memcpy_d2h": (_i, [_p, _p, C.c_size_t]),
    "jlc_dev_synchronize": (_i, []),
    "jlc_dev_copy": (_i, [_p, _p, C.c_size_t, _p]),
    "jlc_stream_get": (_i, [C.POINTER(_p)]),
    "jlc_stream_synchronize": (_i, [_p]),
    "jlc_stream_create": (_i, [C.POINTER(_p)]),
    "jlc_stream_destroy": (_i, [_p]),
    "jlc_timer_create": (_i, [C.POINTER(_p)]),
    "jlc_timer_start": (_i, [_p, _p]),
    "jlc_timer_stop": (_i, [_p, _p, C.POINTER(_d)]),
    "jlc_timer_destroy": (_i, [_p]),
    "jlc_grouped_plan_create": (_i, [_i64, _p, _i, C.POINTER(_p)]),
    "jlc_grouped_plan_run": (_i, [_p, _p, C.POINTER(Stats)]),
    "jlc_grouped_plan_replay": (_i, [_p, _p, C.POINTER(Stats)]),
    "jlc_grouped_plan_destroy": (_i, [_p]),
    "jlc_dgemm_grouped_dev": (_i, [_i64, _p, _i, _p, C.POINTER(Stats)]),
    "jlc_tune_load": (_i, [C.c_char_p]),
    "jlc_tune_write": (_i, [C.c_char_p]),
    "jlc_tune_get": (_i, [C.c_char_p, C.POINTER(_d)]),
    "jlc_tune_set": (_i, [C.c_char_p, _d]),
    "jlc_dgeqrf_dev": (_i, [_i64, _i64, _p, _i64, _p, _i64, _i, _p, C.POINTER(Stats)]),
    "jlc_dorgqr_dev": (_i, [_i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p]),
    "jlc_dqr_errors_dev": (_i, [_i64, _i64, _p, _i64, _p, _i64, _p, _i64, C.POINTER(_d), C.POINTER(_d)]),
    "jlc_dtsqr_dev": (_i, [_i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p, C.POINTER(Stats)]),
    "jlc_dtsqr": (_i, [_i64, _i64, _p, _i64, _p, _i64, _p, _i64, C.POINTER(Stats)]),
    "jlc_qr_batched_plan_create": (_i, [_i64, _p, C.POINTER(_p)]),
    "jlc_qr_batched_plan_replay": (_i, [_p, _p, C.POINTER(Stats)]),
    "jlc_qr_batched_plan_destroy": (_i, [_p]),
    "jlc_dgeqrf_batched_dev": (_i, [_i64, _p, _p, C.POINTER(Stats)]),
}

_lock = threading.Lock()
_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Loads libjlc_hip.so once; raises NativeLibraryMissing if it was not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        table = os.environ.get("JLC_TUNE_TABLE", str(TUNE_TABLE))
        if table and Path(table).exists():
            if lib.jlc_tune_load(table.encode()) != JLC_OK:
                raise JlcError(JLC_ERR_INVALID_ARG, (lib.jlc_last_error() or b"").decode())
        return lib


def last_error() -> str:
    return (load().jlc_last_error() or b"").decode()


def check(rc: int) -> None:
    if rc != JLC_OK:
        raise JlcError(rc, last_error())


def probe() -> tuple[int, int, str]:
    """(status, device_count, name); never raises for a missing device."""
    lib = load()
    cnt = C.c_int(0)
    buf = C.create_string_buffer(256)
    rc = lib.jlc_hip_probe(C.byref(cnt), buf, 256)
    return rc, cnt.value, buf.value.decode()


def _ptr(arr: np.ndarray | None, offset: int = 0) -> int | None:
    if arr is None:
        return None
    if arr.dtype != np.float64:
        raise TypeError("operands must be float64")
    if not arr.flags.c_contiguous:
        raise ValueError("operands must be C-contiguous buffers")
    if offset < 0 or (arr.size and offset > arr.size):
        raise ValueError("offset out of range")
    return arr.ctypes.data + 8 * offset


def dgemm(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a: np.ndarray | None, lda: int,
          b: np.ndarray | None, ldb: int, beta: float, c: np.ndarray, ldc: int, mode: int = MODE_FP64,
          a_off: int = 0, b_off: int = 0, c_off: int = 0) -> Stats:
    """Host-memory GEMM through jlc_dgemm (C updated in place)."""
    st = Stats()
    rc = load().jlc_dgemm(layout_a, layout_b, m, n, k, alpha, _ptr(a, a_off), lda, _ptr(b, b_off), ldb, beta,
                          _ptr(c, c_off), ldc, mode, C.byref(st))
    check(rc)
    return st


LEFT = 0
RIGHT = 1


def wy_apply(side: int, trans_t: int, m: int, n: int, kb: int, v: np.ndarray | None, ldv: int,
             t: np.ndarray | None, ldt: int, c: np.ndarray, ldc: int, mode: int = MODE_FP64, v_off: int = 0,
             t_off: int = 0, c_off: int = 0) -> Stats:
    """Host-memory compact-WY update through jlc_wy_apply (the m x n block of C at c_off updated)."""
    st = Stats()
    check(load().jlc_wy_apply(side, trans_t, m, n, kb, _ptr(v, v_off), ldv, _ptr(t, t_off), ldt, _ptr(c, c_off),
                              ldc, mode, C.byref(st)))
    return st


def wy_apply_dev(side: int, trans_t: int, m: int, n: int, kb: int, v: int | None, ldv: int, t: int | None,
                 ldt: int, c: int, ldc: int, mode: int = MODE_FP64, stream: int | None = None,
                 stats: bool = True) -> Stats | None:
    """Device-pointer compact-WY update through jlc_wy_apply_dev."""
    st = Stats() if stats else None
    check(load().jlc_wy_apply_dev(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, mode, stream,
                                  C.byref(st) if st is not None else None))
    return st


def dgemv(m: int, k: int, alpha: float, a: np.ndarray | None, lda: int, x: np.ndarray | None, incx: int,
          beta: float, y: np.ndarray, incy: int, a_off: int = 0, x_off: int = 0, y_off: int = 0) -> Stats:
    st = Stats()
    rc = load().jlc_dgemv(m, k, alpha, _ptr(a, a_off), lda, _ptr(x, x_off), incx, beta, _ptr(y, y_off), incy,
                          C.byref(st))
    check(rc)
    return st


class DeviceArray:
    """A float64 device allocation owned by the library's allocator."""

    def __init__(self, count: int):
        self.count = int(count)
        p = C.c_void_p()
        check(load().jlc_dev_alloc(C.byref(p), max(1, self.count) * 8))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        d = cls(arr.size)
        if arr.size:
            check(load().jlc_memcpy_h2d(d.ptr, arr.ctypes.data, arr.size * 8))
        return d

    def copy_from(self, arr: np.ndarray) -> None:
        """Overwrites the allocation with arr (same element count)."""
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        if arr.size != self.count:
            raise ValueError(f"{arr.size} elements for a {self.count}-element device array")
        if arr.size:
            check(load().jlc_memcpy_h2d(self.ptr, arr.ctypes.data, arr.size * 8))

    def to_numpy(self, shape=None) -> np.ndarray:
        out = np.empty(self.count, dtype=np.float64)
        if self.count:
            check(load().jlc_memcpy_d2h(out.ctypes.data, self.ptr, self.count * 8))
        return out.reshape(shape) if shape is not None else out

    def at(self, offset: int) -> int:
        return self.ptr + 8 * int(offset)

    def free(self) -> None:
        if self.ptr:
            check(load().jlc_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib is not None:
                _lib.jlc_dev_free(self.ptr)
        except Exception:
            pass


def dgemm_dev(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a: int | None, lda: int,
              b: int | None, ldb: int, beta: float, c: int, ldc: int, mode: int = MODE_FP64, stream: int | None = None,
              stats: bool = True) -> Stats | None:
    """Device-pointer GEMM through jlc_dgemm_dev (pointers are ints)."""
    st = Stats() if stats else None
    rc = load().jlc_dgemm_dev(layout_a, layout_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, mode, stream,
                              C.byref(st) if st is not None else None)
    check(rc)
    return st


def dgemv_dev(m: int, k: int, alpha: float, a: int | None, lda: int, x: int | None, incx: int, beta: float, y: int,
              incy: int, stream: int | None = None, stats: bool = True) -> Stats | None:
    st = Stats() if stats else None
    rc = load().jlc_dgemv_dev(m, k, alpha, a, lda, x, incx, beta, y, incy, stream,
                              C.byref(st) if st is not None else None)
    check(rc)
    return st


def dev_copy(dst: int, src: int, nbytes: int, stream_handle: int | None = None) -> None:
    check(load().jlc_dev_copy(dst, src, nbytes, stream_handle))


def synchronize() -> None:
    check(load().jlc_dev_synchronize())


def stream() -> int:
    p = C.c_void_p()
    check(load().jlc_stream_get(C.byref(p)))
    return p.value


def stream_synchronize(stream_handle: int | None = None) -> None:
    check(load().jlc_stream_synchronize(stream_handle))


def stream_create() -> int:
    p = C.c_void_p()
    check(load().jlc_stream_create(C.byref(p)))
    return p.value


def stream_destroy(stream_handle: int) -> None:
    check(load().jlc_stream_destroy(stream_handle))


class Timer:
    """HIP-event timer on a library stream."""

    def __init__(self, stream_handle: int | None = None):
        p = C.c_void_p()
        check(load().jlc_timer_create(C.byref(p)))
        self.handle = p.value
        self.stream = stream_handle

    def start(self) -> None:
        check(load().jlc_timer_start(self.handle, self.stream))

    def stop(self) -> float:
        ms = C.c_double(0.0)
        check(load().jlc_timer_stop(self.handle, self.stream, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.jlc_timer_destroy(self.handle)
        except Exception:
            pass


class GroupedPlan:
    """A grouped launch over many problems (jlc_grouped_plan_*)."""

    def __init__(self, descs, mode: int = MODE_FP64):
        arr = (GemmDesc * len(descs))(*descs)
        self._arr = arr
        p = C.c_void_p()
        check(load().jlc_grouped_plan_create(len(descs), C.cast(arr, C.c_void_p), mode, C.byref(p)))
        self.handle = p.value

    def run(self, stream_handle: int | None = None, stats: bool = False) -> Stats | None:
        st = Stats() if stats else None
        check(load().jlc_grouped_plan_run(self.handle, stream_handle, C.byref(st) if st is not None else None))
        return st

    def replay(self, stream_handle: int | None = None, stats: bool = False) -> Stats | None:
        st = Stats() if stats else None
        check(load().jlc_grouped_plan_replay(self.handle, stream_handle, C.byref(st) if st is not None else None))
        return st

    def destroy(self) -> None:
        if self.handle:
            check(load().jlc_grouped_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.jlc_grouped_plan_destroy(self.handle)
        except Exception:
            pass


def dgemm_grouped_dev(descs, mode: int = MODE_FP64, stream_handle: int | None = None) -> Stats:
    arr = (GemmDesc * len(descs))(*descs)
    st = Stats()
    check(load().jlc_dgemm_grouped_dev(len(descs), C.cast(arr, C.c_void_p), mode, stream_handle, C.byref(st)))
    return st


def dgeqrf_dev(m: int, n: int, a: int, lda: int, tau: int, nb: int = 128, mode: int = MODE_FP64,
               stream_handle: int | None = None) -> Stats:
    """Blocked Householder QR of the column-major m x n matrix at device pointer a (in place)."""
    st = Stats()
    check(load().jlc_dgeqrf_dev(m, n, a, lda, tau, nb, mode, stream_handle, C.byref(st)))
    return st


def dorgqr_dev(m: int, n: int, a: int, lda: int, tau: int, nb: int, q: int, ldq: int) -> None:
    check(load().jlc_dorgqr_dev(m, n, a, lda, tau, nb, q, ldq, None))


def dqr_errors_dev(m: int, n: int, a_orig: int, lda_orig: int, qr: int, lda: int, tau: int, nb: int) -> tuple:
    """(||A - QR||_F / ||A||_F, ||Q^T Q - I||_F)."""
    r, o = C.c_double(), C.c_double()
    check(load().jlc_dqr_errors_dev(m, n, a_orig, lda_orig, qr, lda, tau, nb, C.byref(r), C.byref(o)))
    return r.value, o.value


def dtsqr_dev(m: int, n: int, a: int, lda: int, q: int, ldq: int, r: int, ldr: int,
              stream_handle: int | None = None) -> Stats:
    """Tall-skinny QR of the column-major device matrix a: thin Q (m x min(m,n)) and R (min(m,n) x n)."""
    st = Stats()
    check(load().jlc_dtsqr_dev(m, n, a, lda, q, ldq, r, ldr, stream_handle, C.byref(st)))
    return st


def dtsqr(m: int, n: int, a: np.ndarray, q: np.ndarray, r: np.ndarray) -> Stats:
    """Tall-skinny QR of the row-major host matrix a (m x n) into row-major q (m x k) and r (k x n)."""
    k = min(m, n)
    st = Stats()
    check(load().jlc_dtsqr(m, n, _ptr(a), max(1, n), _ptr(q), max(1, k), _ptr(r), max(1, n), C.byref(st)))
    return st


class QrBatchedPlan:
    def __init__(self, descs):
        arr = (QrDesc * len(descs))(*descs)
        self._arr = arr
        p = C.c_void_p()
        check(load().jlc_qr_batched_plan_create(len(descs), C.cast(arr, C.c_void_p), C.byref(p)))
        self.handle = p.value

    def replay(self, stream_handle: int | None = None, stats: bool = False) -> Stats | None:
        st = Stats() if stats else None
        check(load().jlc_qr_batched_plan_replay(self.handle, stream_handle, C.byref(st) if st is not None else None))
        return st

    def destroy(self) -> None:
        if self.handle:
            check(load().jlc_qr_batched_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.jlc_qr_batched_plan_destroy(self.handle)
        except Exception:
            pass


def dgeqrf_batched_dev(descs, stream_handle: int | None = None) -> Stats:
    arr = (QrDesc * len(descs))(*descs)
    st = Stats()
    check(load().jlc_dgeqrf_batched_dev(len(descs), C.cast(arr, C.c_void_p), stream_handle, C.byref(st)))
    return st


def tune_get(key: str) -> float:
    v = C.c_double()
    check(load().jlc_tune_get(key.encode(), C.byref(v)))
    return v.value


def tune_set(key: str, value: float) -> None:
    check(load().jlc_tune_set(key.encode(), float(value)))


def tune_write(path: str) -> None:
    check(load().jlc_tune_write(str(path).encode()))


def tune_load(path: str) -> None:
    check(load().jlc_tune_load(str(path).encode()))


def header_symbols() -> list[str]:
    """Function names declared in include/*.h (used by the ABI export test)."""
    import re
    inc = Path(__file__).resolve().parent.parent / "include"
    names: list[str] = []
    for h in sorted(inc.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for mt in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(jlc_\w+)\s*\(", text, flags=re.M):
            names.append(mt.group(1))
    return names


def env_disabled() -> bool:
    v = os.environ.get("JLC_HIP_ENABLED", "")
    return v.lower() in ("0", "false", "off", "no")
