"""Concurrent calls from many host threads: the reference's GEMM entry points are static and are
called from many Java threads at once (OptimizedBLAS3.gemm, src/main/java/net/faulj/compute/
OptimizedBLAS3.java:61, with ThreadLocal workspaces). Each thread owns a stream and workspace in the
library; results must equal the same calls made one at a time, bit for bit (the kernels are
deterministic)."""
import threading

import numpy as np
import pytest

from jlc_amd import native

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(91)
    out = []
    modes = [native.MODE_FP64, native.MODE_SPLIT_CRT, native.MODE_SPLIT_INT8,
             native.mode_slices(native.MODE_SPLIT_CRT, 12)]
    for i in range(24):
        m, n, k = (int(x) for x in rng.integers(1, 300, 3))
        la, lb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        A = rng.uniform(-1, 1, m * k)
        B = rng.uniform(-1, 1, k * n)
        C = rng.uniform(-1, 1, m * n)
        out.append((la, lb, m, n, k, float(rng.uniform(-2, 2)), A, B, float(rng.uniform(-1, 1)), C,
                    modes[i % len(modes)]))
    return out


def _run(case):
    la, lb, m, n, k, alpha, A, B, beta, C, mode = case
    lda = k if la == native.ROW else m
    ldb = n if lb == native.ROW else k
    c = C.copy()
    st = native.dgemm(la, lb, m, n, k, alpha, A, lda, B, ldb, beta, c, n, mode)
    return c, st.kernel_name


def test_many_threads_match_sequential_results():
    cases = _cases()
    expected = [_run(cs) for cs in cases]
    results = [None] * len(cases)
    errors = []

    def worker(tid):
        try:
            for rep in range(3):
                for i in range(tid, len(cases), 8):
                    results[i] = _run(cases[i])
        except Exception as exc:  # reported below
            errors.append(repr(exc))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for i, (got, want) in enumerate(zip(results, expected)):
        assert got is not None, i
        assert got[1] == want[1], (i, got[1], want[1])
        assert np.array_equal(got[0], want[0]), i
