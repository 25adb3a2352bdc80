"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker (or the timed CPU
baseline), never as the product path. See oracle.h for what each function
restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"

ROW = 0
COL = 1

_lib = None
_d = C.c_double
_i64 = C.c_int64
_p = C.c_void_p


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
    lib = C.CDLL(str(LIB_PATH))
    lib.oracle_seeded_array.argtypes = [_p, _i64, _i64]
    lib.oracle_uniform_array.argtypes = [_p, _i64, _i64, _d, _d]
    lib.oracle_splitmix_next.argtypes = [C.POINTER(C.c_uint64)]
    lib.oracle_splitmix_next.restype = C.c_uint64
    lib.oracle_wy_apply.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, C.c_int]
    lib.oracle_gemm_general.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64,
                                        C.c_int]
    for name in ("oracle_gemm_strided", "oracle_gemm_strided_trans_a", "oracle_gemm_strided_col_major_a",
                 "oracle_gemm_strided_col_major_b"):
        getattr(lib, name).argtypes = [_p, C.c_int, C.c_int, _p, C.c_int, C.c_int, _p, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, _d, _d]
    lib.oracle_gemm.argtypes = [_p, _p, _p, C.c_int, C.c_int, C.c_int, _d, _d]
    lib.oracle_parity.argtypes = [_p, _p, _i64, C.POINTER(_d), C.POINTER(_d)]
    lib.oracle_cpu_abs_tol.argtypes = [_p, _i64, _i64]
    lib.oracle_cpu_abs_tol.restype = _d
    lib.oracle_gpu_abs_tol.argtypes = [_p, _i64, _i64]
    lib.oracle_gpu_abs_tol.restype = _d
    lib.oracle_parallel_micro.argtypes = [_i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, C.c_int,
                                          C.POINTER(C.c_int)]
    lib.oracle_block_sizes.argtypes = [C.POINTER(C.c_int)]
    lib.oracle_matvec.argtypes = [_i64, _i64, _d, _p, _i64, _p, _d, _p]
    lib.oracle_dd_entries.argtypes = [C.c_int, C.c_int, _i64, _p, _i64, _p, _i64, _i64, _p, _p, _p, _p, _p,
                                      C.c_int]
    lib.oracle_dd_gemm.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p,
                                   _p, C.c_int]
    lib.oracle_split_exponent.argtypes = [_p, _i64, _i64]
    lib.oracle_split_exponent.restype = C.c_int
    lib.oracle_split_slices.argtypes = [_p, _i64, _i64, C.c_int, C.c_int, _p, _i64]
    lib.oracle_split_gemm.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64,
                                      C.c_int, _i64, C.c_int]
    lib.oracle_split_chunk_k.argtypes = [C.c_int, _i64]
    lib.oracle_split_chunk_k.restype = _i64
    lib.oracle_split_gemm_digits.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p,
                                             _i64, C.c_int, C.c_int, _i64, C.c_int]
    lib.oracle_split_bf16_chunk_k.argtypes = [C.c_int, _i64]
    lib.oracle_split_bf16_chunk_k.restype = _i64
    lib.oracle_crt_modulus.argtypes = [C.c_int]
    lib.oracle_crt_log2_product.argtypes = [C.c_int]
    lib.oracle_crt_log2_product.restype = _d
    lib.oracle_crt_tables.argtypes = [C.c_int, _p, _p, _p]
    lib.oracle_crt_norm_exponent.argtypes = [_p, _i64, _i64, C.c_int]
    lib.oracle_crt_norm_exponent.restype = C.c_int
    lib.oracle_crt_residues.argtypes = [_p, _i64, _i64, C.c_int, C.c_int, C.c_int, _p, _i64]
    lib.oracle_crt_reconstruct.argtypes = [_p, C.c_int, _p, _p, _p]
    lib.oracle_crt_reconstruct.restype = _d
    lib.oracle_crt_gemm.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64,
                                    C.c_int, C.c_int, C.c_int, C.c_int]
    lib.oracle_qr_factor.argtypes = [_i64, _i64, _p, _i64, _p, _i64]
    lib.oracle_qr_form_q.argtypes = [_i64, _i64, _p, _i64, _p, _p, _i64]
    lib.oracle_tsqr.argtypes = [_i64, _i64, _p, _i64, _p, _i64, _p, _i64]
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None, off: int = 0):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data + 8 * off


def threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def seeded_array(length: int, seed: int) -> np.ndarray:
    """GemmReference.seededArray (GemmReference.java:14-25)."""
    out = np.empty(length, dtype=np.float64)
    _load().oracle_seeded_array(_ptr(out), length, seed)
    return out


def uniform_array(length: int, seed: int, origin: float = -1.0, bound: float = 1.0) -> np.ndarray:
    out = np.empty(length, dtype=np.float64)
    _load().oracle_uniform_array(_ptr(out), length, seed, origin, bound)
    return out


def splitmix_stream(seed: int, count: int) -> list[int]:
    st = C.c_uint64(seed & (2**64 - 1))
    lib = _load()
    return [lib.oracle_splitmix_next(C.byref(st)) for _ in range(count)]


def gemm_general(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int,
                 beta: float, c: np.ndarray, ldc: int, a_off: int = 0, b_off: int = 0, c_off: int = 0,
                 nthreads: int | None = None) -> None:
    """In-place c <- alpha*op(A)op(B) + beta*c (GemmReference.java:31-112 index maps)."""
    _load().oracle_gemm_general(layout_a, layout_b, m, n, k, alpha, _ptr(a, a_off), lda, _ptr(b, b_off), ldb, beta,
                                _ptr(c, c_off), ldc, nthreads or threads())


def gemm(a, b, c, m, k, n, alpha, beta) -> np.ndarray:
    """GemmReference.gemm: returns a new array (GemmReference.java:31-44)."""
    out = np.array(c, dtype=np.float64, copy=True)
    gemm_general(ROW, ROW, m, n, k, alpha, a, max(1, k), b, max(1, n), beta, out, max(1, n))
    return out


def extract_row_major(backing: np.ndarray, offset: int, ld: int, rows: int, cols: int) -> np.ndarray:
    """GemmReference.extractRowMajor (GemmReference.java:114-120)."""
    out = np.empty(rows * cols, dtype=np.float64)
    for i in range(rows):
        out[i * cols:(i + 1) * cols] = backing[offset + i * ld: offset + i * ld + cols]
    return out


def to_column_major(row_major: np.ndarray, rows: int, cols: int, ld: int) -> np.ndarray:
    """GemmReference.toColumnMajor (GemmReference.java:122-130)."""
    out = np.zeros(max(0, cols * ld), dtype=np.float64)
    for i in range(rows):
        for j in range(cols):
            out[j * ld + i] = row_major[i * cols + j]
    return out


def _strided(kind: str, a, aOff, lda, b, bOff, ldb, c, cOff, ldc, m, k, n, alpha, beta) -> np.ndarray:
    work = np.array(c, dtype=np.float64, copy=True)
    lib = _load()
    getattr(lib, kind)(_ptr(a), aOff, lda, _ptr(b), bOff, ldb, _ptr(work), cOff, ldc, m, k, n, alpha, beta)
    rows = k if kind == "oracle_gemm_strided_trans_a" else m
    return extract_row_major(work, cOff, ldc, rows, n)


def gemm_strided(*args) -> np.ndarray:
    """GemmReference.gemmStrided (GemmReference.java:46-61); returns the row-major C block."""
    return _strided("oracle_gemm_strided", *args)


def gemm_strided_trans_a(*args) -> np.ndarray:
    """GemmReference.gemmStridedTransA (GemmReference.java:63-78); C is k x n."""
    return _strided("oracle_gemm_strided_trans_a", *args)


def gemm_strided_col_major_a(*args) -> np.ndarray:
    """GemmReference.gemmStridedColMajorA (GemmReference.java:80-95)."""
    return _strided("oracle_gemm_strided_col_major_a", *args)


def gemm_strided_col_major_b(*args) -> np.ndarray:
    """GemmReference.gemmStridedColMajorB (GemmReference.java:97-112)."""
    return _strided("oracle_gemm_strided_col_major_b", *args)


def parity(expected: np.ndarray, actual: np.ndarray) -> tuple[float, float]:
    """(maxAbsErr, relErr) as in GemmReference.assertParity (GemmReference.java:148-173)."""
    e = np.ascontiguousarray(expected, dtype=np.float64).reshape(-1)
    a = np.ascontiguousarray(actual, dtype=np.float64).reshape(-1)
    assert e.size == a.size
    mx, rel = _d(0.0), _d(0.0)
    _load().oracle_parity(_ptr(e), _ptr(a), e.size, C.byref(mx), C.byref(rel))
    return mx.value, rel.value


def cpu_abs_tolerance(expected: np.ndarray, k: int) -> float:
    """GemmReference.cpuAbsTolerance (GemmReference.java:175-177)."""
    e = np.ascontiguousarray(expected, dtype=np.float64).reshape(-1)
    return _load().oracle_cpu_abs_tol(_ptr(e), e.size, k)


def gpu_abs_tolerance(expected: np.ndarray, k: int) -> float:
    """GemmReference.cudaAbsTolerance (GemmReference.java:179-181)."""
    e = np.ascontiguousarray(expected, dtype=np.float64).reshape(-1)
    return _load().oracle_gpu_abs_tol(_ptr(e), e.size, k)


def assert_parity(path: str, m: int, k: int, n: int, alpha: float, beta: float, seed: int, expected, actual,
                  rel_tol: float, abs_tol: float) -> None:
    """GemmReference.assertParity (GemmReference.java:148-173)."""
    max_abs, rel = parity(expected, actual)
    ref_norm = float(np.linalg.norm(np.asarray(expected, dtype=np.float64)))
    msg = (f"{path} parity failed for shape m={m} k={k} n={n} alpha={alpha} beta={beta} seed={seed} "
           f"maxAbsErr={max_abs} relErr={rel} absTol={abs_tol} relTol={rel_tol}")
    assert max_abs <= abs_tol, msg
    if ref_norm != 0.0:
        assert rel <= rel_tol, msg


def parallel_micro(m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int, beta: float, c: np.ndarray,
                   ldc: int, max_threads: int | None = None) -> int:
    """PARALLEL_MICRO restatement (OptimizedBLAS3.java:295-385); returns threads used."""
    used = C.c_int(0)
    _load().oracle_parallel_micro(m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc,
                                  max_threads or threads(), C.byref(used))
    return used.value


def block_sizes() -> dict:
    v = (C.c_int * 5)()
    _load().oracle_block_sizes(v)
    return dict(mc=v[0], kc=v[1], nc=v[2], mr=v[3], nr=v[4])


def matvec(m: int, k: int, alpha: float, a, lda: int, x, beta: float, y: np.ndarray) -> None:
    """SpecializedKernels.matvec (SpecializedKernels.java:27-71), y in place."""
    _load().oracle_matvec(m, k, alpha, _ptr(a), lda, _ptr(x), beta, _ptr(y))


def dd_entries(layout_a: int, layout_b: int, k: int, a, lda: int, b, ldb: int, ii, jj,
               nthreads: int | None = None):
    """Double-double sums of op(A)(i,:)·op(B)(:,j) for each (i,j): (hi, lo, sum|a||b|)."""
    ii = np.ascontiguousarray(ii, dtype=np.int64)
    jj = np.ascontiguousarray(jj, dtype=np.int64)
    cnt = ii.size
    hi = np.empty(cnt)
    lo = np.empty(cnt)
    ab = np.empty(cnt)
    _load().oracle_dd_entries(layout_a, layout_b, k, _ptr(a), lda, _ptr(b), ldb, cnt, ii.ctypes.data,
                              jj.ctypes.data, _ptr(hi), _ptr(lo), _ptr(ab), nthreads or threads())
    return hi, lo, ab


def dd_gemm(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int,
            beta: float, c, ldc: int, nthreads: int | None = None):
    hi = np.empty(m * n)
    lo = np.empty(m * n)
    _load().oracle_dd_gemm(layout_a, layout_b, m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc,
                           _ptr(hi), _ptr(lo), nthreads or threads())
    return hi, lo


def wy_apply(side: int, trans_t: int, m: int, n: int, kb: int, v, ldv: int, t, ldt: int, c: np.ndarray, ldc: int,
             c_off: int = 0, nthreads: int | None = None) -> None:
    """Compact-WY block update of C in place (applyBlockLeft / applyBlockRight restated; oracle.c)."""
    _load().oracle_wy_apply(side, trans_t, m, n, kb, _ptr(v), ldv, _ptr(t), ldt, _ptr(c, c_off), ldc,
                            nthreads or threads())


def qr_factor(a_colmajor: np.ndarray, m: int, n: int, lda: int | None = None, nb: int = 64):
    """HouseholderQR restatement: factors in place (column-major), returns tau."""
    lda = lda or m
    tau = np.zeros(min(m, n))
    _load().oracle_qr_factor(m, n, _ptr(a_colmajor), lda, _ptr(tau), nb)
    return tau


def tsqr(a_colmajor: np.ndarray, m: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    """TallSkinnyQR.decompose restated: (Q column-major m x k, R column-major k x n), k = min(m, n)."""
    k = min(m, n)
    q = np.zeros(m * k)
    r = np.zeros(k * n)
    _load().oracle_tsqr(m, n, _ptr(np.ascontiguousarray(a_colmajor, dtype=np.float64)), m, _ptr(q), m, _ptr(r), k)
    return q, r


def qr_form_q(qr_colmajor: np.ndarray, tau: np.ndarray, m: int, n: int, lda: int | None = None) -> np.ndarray:
    lda = lda or m
    q = np.zeros(m * n)
    _load().oracle_qr_form_q(m, n, _ptr(qr_colmajor), lda, _ptr(tau), _ptr(q), m)
    return q


SPLIT_NONFINITE = -100000


def split_exponent(x: np.ndarray) -> int:
    """Row / column exponent of the split form: frexp exponent of max |x|."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _load().oracle_split_exponent(_ptr(x), 1, x.size)


def split_slices(x: np.ndarray, e: int, slices: int) -> np.ndarray:
    """(slices, len(x)) int8 digits of x * 2^-e."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((slices, x.size), dtype=np.int8)
    _load().oracle_split_slices(_ptr(x), 1, x.size, e, slices, out.ctypes.data, x.size)
    return out


def split_chunk_k(slices: int, k: int) -> int:
    return _load().oracle_split_chunk_k(slices, k)


def split_gemm(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int,
               beta: float, c: np.ndarray, ldc: int, slices: int, chunk_k: int | None = None, a_off: int = 0,
               b_off: int = 0, c_off: int = 0, nthreads: int | None = None) -> None:
    """In-place split-precision GEMM restatement (bit-exact reference of the INT8 split form)."""
    ck = chunk_k if chunk_k is not None else split_chunk_k(slices, k)
    _load().oracle_split_gemm(layout_a, layout_b, m, n, k, alpha, _ptr(a, a_off), lda, _ptr(b, b_off), ldb, beta,
                              _ptr(c, c_off), ldc, slices, ck, nthreads or threads())


def split_bf16_chunk_k(slices: int, k: int) -> int:
    return _load().oracle_split_bf16_chunk_k(slices, k)


CRT_MAX_MODULI = 20


def crt_moduli(nmod: int = CRT_MAX_MODULI) -> list[int]:
    lib = _load()
    return [lib.oracle_crt_modulus(t) for t in range(nmod)]


def crt_log2_product(nmod: int) -> float:
    return _load().oracle_crt_log2_product(nmod)


def crt_tables(nmod: int):
    """(w[nmod, 4], mlimb[4], finv[4]) of the CRT reconstruction."""
    w = np.zeros(nmod * 4)
    ml = np.zeros(4)
    fi = np.zeros(4)
    _load().oracle_crt_tables(nmod, _ptr(w), _ptr(ml), _ptr(fi))
    return w.reshape(nmod, 4), ml, fi


def crt_residues(x: np.ndarray, e: int, p: int, nmod: int) -> np.ndarray:
    """Symmetric residues (nmod x len(x), int8) of the p-bit integer images of x (row exponent e)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((nmod, x.size), dtype=np.int8)
    _load().oracle_crt_residues(_ptr(x), 1, x.size, e, p, nmod, out.ctypes.data, x.size)
    return out


def crt_reconstruct(residues, nmod: int) -> float:
    w, ml, fi = crt_tables(nmod)
    r = np.ascontiguousarray(residues, dtype=np.int32)
    return _load().oracle_crt_reconstruct(r.ctypes.data, nmod, _ptr(w), _ptr(ml), _ptr(fi))


def crt_gemm(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int,
             beta: float, c, ldc: int, nmod: int, pa: int, pb: int, a_off: int = 0, b_off: int = 0, c_off: int = 0,
             nthreads: int | None = None) -> None:
    """In-place CRT split-form GEMM restatement (bit-exact reference of the CRT form)."""
    _load().oracle_crt_gemm(layout_a, layout_b, m, n, k, alpha, _ptr(a, a_off), lda, _ptr(b, b_off), ldb, beta,
                            _ptr(c, c_off), ldc, nmod, pa, pb, nthreads or threads())


CRT_MAX_BITS = 65


def crt_norm_exponent(x: np.ndarray, e: int | None = None) -> int:
    """t with ||x||_2 <= 2^t from the integer norm bound (e: max exponent; computed when None)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib = _load()
    if e is None:
        e = lib.oracle_split_exponent(_ptr(x), 1, x.size)
    return lib.oracle_crt_norm_exponent(_ptr(x), 1, x.size, e)


def crt_fits(pa: int, pb: int, k: int, nmod: int) -> bool:
    """Whether nmod moduli hold C' for norm bits (pa, pb): (2^pa + sqrt(k)/2)(2^pb + sqrt(k)/2) * 4 <= M."""
    h = 0.5 * np.sqrt(float(k))
    need = np.log2(np.ldexp(1.0, pa) + h) + np.log2(np.ldexp(1.0, pb) + h) + 2.0
    return bool(need <= crt_log2_product(nmod) - 1e-6)


def crt_bits_for(nmod: int, k: int) -> tuple[int, int]:
    """(pa, pb) the CRT form uses for an explicit moduli count: the largest even
    split of the norm bits that the moduli hold, at most 65 each."""
    for total in range(2 * CRT_MAX_BITS, 3, -1):
        pa = total // 2
        if crt_fits(pa, total - pa, k, nmod):
            return pa, total - pa
    raise ValueError(f"{nmod} moduli leave no bits for k={k}")


def split_gemm_bf16(layout_a: int, layout_b: int, m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int,
                    beta: float, c: np.ndarray, ldc: int, slices: int, chunk_k: int | None = None, a_off: int = 0,
                    b_off: int = 0, c_off: int = 0, nthreads: int | None = None) -> None:
    """In-place restatement of the BF16-slice form (8-bit digits)."""
    ck = chunk_k if chunk_k is not None else split_bf16_chunk_k(slices, k)
    _load().oracle_split_gemm_digits(layout_a, layout_b, m, n, k, alpha, _ptr(a, a_off), lda, _ptr(b, b_off), ldb,
                                     beta, _ptr(c, c_off), ldc, slices, 8, ck, nthreads or threads())
