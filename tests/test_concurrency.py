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


def test_one_thread_two_streams_match_sequential_results():
    """One thread's device-pointer calls on two streams share the thread's split-form and split-K
    workspace (residue planes, slices, partial sums). With explicit moduli / slice counts the calls are
    asynchronous, so the second stream's call must wait for the first stream's kernels before it
    overwrites that workspace. Results must equal the same calls made one at a time, bit for bit."""
    rng = np.random.default_rng(17)
    crt, i8 = native.mode_slices(native.MODE_SPLIT_CRT, 12), native.mode_slices(native.MODE_SPLIT_INT8, 5)
    # (m, n, k, mode): the first case of each pair sizes the workspace, the second reuses it smaller
    shapes = [(1536, 1536, 1536, crt), (1024, 1024, 1024, crt), (1280, 1280, 1280, i8), (768, 896, 640, i8),
              (128, 128, 65536, native.MODE_FP64), (96, 112, 40000, native.MODE_FP64)]
    cases = []
    for m, n, k, mode in shapes:
        A = rng.uniform(-1, 1, m * k)
        B = rng.uniform(-1, 1, k * n)
        cases.append((m, n, k, mode, native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(B),
                      native.DeviceArray(m * n)))

    def call(case, stream, stats):
        m, n, k, mode, da, db, dc = case
        return native.dgemm_dev(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, dc.ptr, n, mode,
                                stream, stats=stats)

    expected = []
    for case in cases:  # one at a time, largest of each pair first
        st = call(case, None, True)
        expected.append((case[6].to_numpy(), st.kernel_name))
    assert [e[1] for e in expected] == ["split_crt", "split_crt", "split_int8", "split_int8", "mfma_f64",
                                        "mfma_f64"]
    s1, s2 = native.stream_create(), native.stream_create()
    try:
        # rep 0, 1: two caller streams (the workspace's end is marked on the stream at each call's end);
        # rep 2: the thread's own stream (NULL), whose last use is marked only when the call on stream 2
        # arrives; rep 3: stream 2 first, then the thread's own stream
        for rep, (first, second) in enumerate(((s1, s2), (s1, s2), (None, s2), (s2, None))):
            for case in cases:  # clear C so stale results cannot pass
                native.DeviceArray.copy_from(case[6], np.full(case[0] * case[1], np.nan))
            for i in range(0, len(cases), 2):
                call(cases[i], first, False)        # asynchronous
                call(cases[i + 1], second, False)   # same thread, same workspace, the other stream
            native.synchronize()
            for i, case in enumerate(cases):
                assert np.array_equal(case[6].to_numpy(), expected[i][0]), (rep, i, case[:4])
    finally:
        native.stream_destroy(s1)
        native.stream_destroy(s2)
