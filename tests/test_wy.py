"""Compact-WY block updates: the three-GEMM sequences of BlockedHessenbergQR.applyBlockLeft /
applyBlockRight (src/main/java/net/faulj/eigen/qr/BlockedHessenbergQR.java:277-345) and of the left / right
block updates of BlockedBidiagonalization (src/main/java/net/faulj/decomposition/bidiagonal/
BlockedBidiagonalization.java:379-401, :581-603), one jlc_wy_apply[_dev] call each.

The oracle's restatement (oracle_wy_apply) is pinned on the CPU by the identity the update exists for:
with T built by the reference's recurrence (buildT, BlockedHessenbergQR.java:239-274),
I - V T V^T = H_0 H_1 ... H_{kb-1} for the reflectors H_j = I - tau_j v_j v_j^T, so each side and
transpose choice equals a product of single reflectors applied one at a time (a transposed T or a swapped
side fails it), and the update preserves Frobenius norms. The GPU tests compare the library with the
restatement at n = 4096 with the reference's panel widths (32 for Hessenberg, 64 for bidiagonalization),
at the reference's CPU tolerance (relTol 1e-12, GemmReference.java:175-177).
"""
import zlib

import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc


def reflectors(rng, rows: int, kb: int):
    """Unit lower-trapezoidal V (rows x kb), tau_j = 2 / v_j^T v_j, and T by buildT's recurrence:
    T[i,i] = tau_i, T[:i, i] = T[:i, :i] w with w_l = -tau_i v_l^T v_i (BlockedHessenbergQR.java:239-274)."""
    V = np.tril(0.3 * rng.standard_normal((rows, kb)), -1)
    V[np.arange(kb), np.arange(kb)] = 1.0
    tau = 2.0 / np.sum(V * V, axis=0)
    T = np.zeros((kb, kb))
    for i in range(kb):
        T[i, i] = tau[i]
        if i:
            T[:i, i] = T[:i, :i] @ (-tau[i] * (V[:, :i].T @ V[:, i]))
    return V, T, tau


def product_of_reflectors(V, tau, reverse: bool) -> np.ndarray:
    rows, kb = V.shape
    H = np.eye(rows)
    for j in (reversed(range(kb)) if reverse else range(kb)):
        v = V[:, j]
        H = H @ (np.eye(rows) - tau[j] * np.outer(v, v))
    return H


@pytest.mark.parametrize("side,trans_t", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_oracle_wy_apply_equals_product_of_single_reflectors(side, trans_t):
    rng = np.random.default_rng(31 + 2 * side + trans_t)
    m, n, kb = 40, 33, 7
    rows = m if side == 0 else n
    V, T, tau = reflectors(rng, rows, kb)
    C = rng.uniform(-1, 1, (m, n))
    got = C.reshape(-1).copy()
    orc.wy_apply(side, trans_t, m, n, kb, np.ascontiguousarray(V).reshape(-1), kb, T.reshape(-1).copy(), kb, got, n)
    got = got.reshape(m, n)
    # I - V T V^T = H_0 ... H_{kb-1}; transposing T reverses the order (each H_j is symmetric)
    H = product_of_reflectors(V, tau, reverse=bool(trans_t))
    want = H @ C if side == 0 else C @ H
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(C))
    assert abs(np.linalg.norm(got) - np.linalg.norm(C)) <= 1e-13 * np.linalg.norm(C)


def test_oracle_wy_apply_zero_panel_is_identity():
    C = np.arange(12.0)
    got = C.copy()
    orc.wy_apply(0, 0, 3, 4, 0, np.zeros(1), 1, np.zeros(1), 1, got, 4)
    assert np.array_equal(got, C)


# ------------------------------------------------------------------------------------------ GPU

def _parity(name, expected, actual, k):
    orc.assert_parity(name, expected.size, k, 1, 1.0, 1.0, 0, expected, actual, 1e-12,
                      orc.cpu_abs_tolerance(expected, k))


def _device_case(side, trans_t, n_full, r0, c0, kb, seed, mode=native.MODE_FP64):
    """The block a[r0:, c0:] of a row-major n_full x n_full matrix updated by jlc_wy_apply_dev."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, n_full * n_full)
    m, n = n_full - r0, n_full - c0
    V, T, _ = reflectors(rng, m if side == native.LEFT else n, kb)
    Vf, Tf = np.ascontiguousarray(V).reshape(-1), T.reshape(-1).copy()
    dA, dV, dT = (native.DeviceArray.from_numpy(x) for x in (A, Vf, Tf))
    st = native.wy_apply_dev(side, trans_t, m, n, kb, dV.ptr, kb, dT.ptr, kb, dA.at(r0 * n_full + c0), n_full, mode)
    expected = A.copy()
    orc.wy_apply(side, trans_t, m, n, kb, Vf, kb, Tf, kb, expected, n_full, c_off=r0 * n_full + c0)
    return expected, dA.to_numpy(), max(m, n), st


@pytest.mark.gpu
@pytest.mark.parametrize("side,trans_t,r0,c0,kb,what", [
    (native.LEFT, 1, 1, 0, 32, "Hessenberg applyBlockLeft, k = 0"),
    (native.RIGHT, 0, 0, 1, 32, "Hessenberg applyBlockRight, k = 0"),
    (native.LEFT, 1, 2017, 2016, 32, "Hessenberg applyBlockLeft, k = 2016"),
    (native.RIGHT, 0, 0, 2017, 32, "Hessenberg applyBlockRight, k = 2016"),
    (native.LEFT, 0, 64, 64, 64, "bidiagonal left update, panel 64"),
    (native.RIGHT, 0, 64, 128, 64, "bidiagonal right update, panel 64"),
])
def test_device_wy_apply_n4096_against_restatement(side, trans_t, r0, c0, kb, what):
    expected, got, k, st = _device_case(side, trans_t, 4096, r0, c0, kb, seed=zlib.crc32(what.encode()))
    assert st.kernel_name == "wy"
    _parity(what, expected, got, k)


def _poisoned_case(side, trans_t, m, n, kb, seed, mode=native.MODE_FP64, col_off=0, ldc=None):
    """An m x n block at column col_off of a larger row-major array (ldc, default n + 3; one extra row)
    whose surroundings, V's columns past kb (ldv = kb + 2) and T's past kb (ldt = kb + 1) hold Inf: none
    of it is part of the operation, so the update must stay finite, leave the surroundings alone and
    match the restatement."""
    rng = np.random.default_rng(seed)
    ldc = ldc or n + col_off + 3
    ldv, ldt = kb + 2, kb + 1
    full = np.full((m + 1, ldc), np.inf)
    full[:m, col_off:col_off + n] = rng.uniform(-1.0, 1.0, (m, n))
    rows = m if side == native.LEFT else n
    V, T, _ = reflectors(rng, rows, kb) if rows >= kb else (0.3 * rng.standard_normal((rows, kb)),
                                                          0.1 * np.triu(rng.standard_normal((kb, kb))), None)
    Vp = np.full((rows, ldv), np.inf)
    Vp[:, :kb] = V
    Tp = np.full((kb, ldt), np.inf)
    Tp[:, :kb] = T
    Vf, Tf, Cf = Vp.reshape(-1).copy(), Tp.reshape(-1).copy(), full.reshape(-1).copy()
    dV, dT, dC = (native.DeviceArray.from_numpy(x) for x in (Vf, Tf, Cf))
    st = native.wy_apply_dev(side, trans_t, m, n, kb, dV.ptr, ldv, dT.ptr, ldt, dC.at(col_off), ldc, mode)
    expected = Cf.copy()
    orc.wy_apply(side, trans_t, m, n, kb, Vf, ldv, Tf, ldt, expected, ldc, c_off=col_off)
    got = dC.to_numpy()
    inside = np.zeros((m + 1, ldc), bool)
    inside[:m, col_off:col_off + n] = True
    inside = inside.reshape(-1)
    assert np.array_equal(got[~inside], expected[~inside]), "entries outside the block changed"
    assert np.all(np.isfinite(got[inside])), "a value outside the operation reached the block"
    return expected[inside], got[inside], max(m, n, kb), st


# Ragged widths (kb not a multiple of 16, one reflector, the 64 limit), blocks smaller than one tile,
# one row or column, heights that leave a partial slice or a partial unrolled row group.
WY_SHAPES = [(1, 1, 1), (5, 3, 2), (63, 65, 7), (65, 63, 16), (129, 300, 30), (300, 129, 33), (1000, 17, 48),
             (17, 1000, 64), (777, 555, 64), (2049, 1031, 32), (1, 700, 5), (700, 1, 5)]


@pytest.mark.gpu
@pytest.mark.parametrize("side", [native.LEFT, native.RIGHT])
@pytest.mark.parametrize("trans_t", [0, 1])
def test_wy_kernels_ragged_shapes_against_restatement(side, trans_t):
    for m, n, kb in WY_SHAPES:
        what = f"side {side} trans {trans_t} m {m} n {n} kb {kb}"
        print(what, flush=True)
        expected, got, k, st = _poisoned_case(side, trans_t, m, n, kb, seed=zlib.crc32(what.encode()))
        assert st.kernel_name == "wy", what
        _parity(what, expected, got, k)


@pytest.mark.gpu
@pytest.mark.parametrize("side", [native.LEFT, native.RIGHT])
def test_wy_kernels_line_shifted_tiles_against_restatement(side):
    """Blocks whose first column is not on a 128-byte line of a row stride that is a multiple of 16
    doubles (H[:, k+1:] of the Hessenberg driver): C's tiles then start before the block, and the columns
    before it must neither be read into a sum nor written."""
    for col_off in (1, 7, 15, 16, 33):
        for m, n, kb in ((300, 200, 32), (129, 1000, 64), (1000, 70, 30), (64, 64, 16)):
            ldc = ((n + col_off + 5 + 15) // 16) * 16
            what = f"side {side} offset {col_off} m {m} n {n} kb {kb} ldc {ldc}"
            print(what, flush=True)
            expected, got, k, st = _poisoned_case(side, side, m, n, kb, seed=zlib.crc32(what.encode()),
                                                  col_off=col_off, ldc=ldc)
            assert st.kernel_name == "wy", what
            _parity(what, expected, got, k)


@pytest.mark.gpu
def test_wy_kernels_slice_counts_and_gemm_sequence_agree():
    """The partial-W slice count (wy_left/right_wgs_per_cu 1, 2, 8: one to many slots, a last slot shorter
    than the rest) and the three-GEMM sequence (0) all match the restatement; kb = 65 takes the GEMM
    sequence by itself."""
    keys = ("wy_left_wgs_per_cu", "wy_right_wgs_per_cu")
    saved = [native.tune_get(k) for k in keys]
    try:
        for wgs in (1, 2, 8, 0):
            for k in keys:
                native.tune_set(k, wgs)
            for side, trans_t, m, n, kb in ((native.LEFT, 1, 3001, 2500, 32), (native.RIGHT, 0, 2500, 3001, 64)):
                what = f"wgs {wgs} side {side} m {m} n {n} kb {kb}"
                print(what, flush=True)
                expected, got, k, st = _poisoned_case(side, trans_t, m, n, kb, seed=zlib.crc32(what.encode()))
                assert st.kernel_name == ("wy" if wgs else "mfma_f64"), what
                _parity(what, expected, got, k)
        for k, v in zip(keys, saved):
            native.tune_set(k, v)
        expected, got, k, st = _poisoned_case(native.LEFT, 0, 300, 200, 65, seed=65)
        assert st.kernel_name == "mfma_f64"
        _parity("kb 65", expected, got, k)
    finally:
        for k, v in zip(keys, saved):
            native.tune_set(k, v)


@pytest.mark.gpu
def test_device_wy_apply_split_crt_mode():
    expected, got, k, st = _device_case(native.LEFT, 0, 2048, 64, 64, 64, seed=77, mode=native.MODE_SPLIT_CRT)
    assert st.kernel_name == "split_crt"
    orc.assert_parity("WY left, CRT products", expected.size, k, 1, 1.0, 1.0, 0, expected, got, 1e-11,
                      orc.cpu_abs_tolerance(expected, k))


@pytest.mark.gpu
def test_hessenberg_mirror_host_path_against_restatement():
    """BlockedHessenbergQR.applyBlockLeft / applyBlockRight through the host mirror (jlc_wy_apply on heap
    arrays), the driver's call pattern for one panel (BlockedHessenbergQR.java:121-133), and the general
    applyBlockRight form with colStart > k + 1."""
    from jlc_amd.compute import BlockedHessenbergQR
    rng = np.random.default_rng(512)
    n, k, pw = 512, 64, 32
    m = n - (k + 1)
    V, T, _ = reflectors(rng, m, pw)
    Vf, Tf = np.ascontiguousarray(V).reshape(-1), T.reshape(-1).copy()
    h = rng.uniform(-1, 1, n * n)
    q = np.eye(n).reshape(-1).copy()
    want_h, want_q = h.copy(), q.copy()
    BlockedHessenbergQR.applyBlockLeft(h, n, k, k, Vf, Tf, m, pw)
    BlockedHessenbergQR.applyBlockRight(h, n, k, k + 1, Vf, Tf, m, pw)
    BlockedHessenbergQR.applyBlockRight(q, n, k, k + 1, Vf, Tf, m, pw)
    orc.wy_apply(0, 1, m, n - k, pw, Vf, pw, Tf, pw, want_h, n, c_off=(k + 1) * n + k)
    orc.wy_apply(1, 0, n, m, pw, Vf, pw, Tf, pw, want_h, n, c_off=k + 1)
    orc.wy_apply(1, 0, n, m, pw, Vf, pw, Tf, pw, want_q, n, c_off=k + 1)
    _parity("applyBlockLeft+Right on H", want_h, h, n)
    _parity("applyBlockRight on Q", want_q, q, n)
    # colStart = k + 4: W from all trailing columns, the update on the last n - colStart of them
    a = rng.uniform(-1, 1, n * n)
    want = a.reshape(n, n).copy()
    BlockedHessenbergQR.applyBlockRight(a, n, k, k + 4, Vf, Tf, m, pw)
    W2 = (want[:, k + 1:] @ V) @ T
    want[:, k + 4:] -= W2 @ V[3:].T
    _parity("applyBlockRight colStart = k + 4", want.reshape(-1), a, n)
    # the reference's early returns leave the array untouched
    before = a.copy()
    BlockedHessenbergQR.applyBlockLeft(a, n, k, n, Vf, Tf, m, pw)
    BlockedHessenbergQR.applyBlockRight(a, n, n - 1, n, Vf, Tf, m, pw)
    assert np.array_equal(a, before)


@pytest.mark.gpu
def test_wy_apply_arguments():
    d = native.DeviceArray(64)
    with pytest.raises(native.JlcError):
        native.wy_apply_dev(2, 0, 4, 4, 2, d.ptr, 2, d.ptr, 2, d.ptr, 4)
    with pytest.raises(native.JlcError):
        native.wy_apply_dev(native.LEFT, 0, 4, 4, 2, d.ptr, 1, d.ptr, 2, d.ptr, 4)  # ldv < kb
    c = np.arange(16.0)
    native.wy_apply(native.LEFT, 0, 4, 4, 0, None, 1, None, 1, c, 4)  # kb == 0: unchanged
    assert np.array_equal(c, np.arange(16.0))


@pytest.mark.gpu
def test_wy_kernels_on_two_streams_match_sequential_results():
    """Two asynchronous WY updates from one thread share the thread's WY workspace (partial slots and W2):
    the first on the thread's own stream, whose end is marked only when the second call arrives on a
    caller's stream, then the reverse. Results equal the same calls made one at a time, bit for bit."""
    rng = np.random.default_rng(4242)
    cases = []
    for side, m, n, kb in ((native.LEFT, 3000, 2000, 32), (native.RIGHT, 1500, 2500, 64),
                           (native.RIGHT, 2200, 1800, 48), (native.LEFT, 900, 1300, 16)):
        rows = m if side == native.LEFT else n
        V, T, _ = reflectors(rng, rows, kb)
        C0 = rng.uniform(-1.0, 1.0, m * n)
        cases.append((side, m, n, kb, native.DeviceArray.from_numpy(np.ascontiguousarray(V).reshape(-1)),
                      native.DeviceArray.from_numpy(T.reshape(-1).copy()), C0, native.DeviceArray(m * n)))

    def call(case, stream):
        side, m, n, kb, dv, dt, _, dc = case
        native.wy_apply_dev(side, 0, m, n, kb, dv.ptr, kb, dt.ptr, kb, dc.ptr, n, native.MODE_FP64, stream,
                            stats=False)

    expected = []
    for case in cases:
        native.DeviceArray.copy_from(case[7], case[6])
        call(case, None)
        native.synchronize()
        expected.append(case[7].to_numpy())
    s2 = native.stream_create()
    try:
        for first, second in ((None, s2), (s2, None)):
            for case in cases:
                native.DeviceArray.copy_from(case[7], case[6])
            native.synchronize()
            for i in range(0, len(cases), 2):
                call(cases[i], first)
                call(cases[i + 1], second)
            native.synchronize()
            for i, case in enumerate(cases):
                assert np.array_equal(case[7].to_numpy(), expected[i]), (first, i, case[:4])
    finally:
        native.stream_destroy(s2)
