"""jlc_dgemm_dev_noalloc, the entry point that allocates nothing: its results against jlc_dgemm_dev's
(the same FP64 kernels, so bit for bit) on two calls in a row, the second smaller, with shapes that
put split-K partial products in the library's fixed device scratch; and the clear error for a shape
whose partials would not fit it."""
import ctypes as C

import numpy as np
import pytest

from jlc_amd import native

pytestmark = pytest.mark.gpu


def _noalloc(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    return native.load().jlc_dgemm_dev_noalloc(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream)


def _case(rng, la, lb, m, n, k):
    A = rng.uniform(-1, 1, m * k)
    B = rng.uniform(-1, 1, k * n)
    C0 = rng.uniform(-1, 1, m * n)
    lda = k if la == native.ROW else m
    ldb = n if lb == native.ROW else k
    return A, B, C0, lda, ldb


def test_noalloc_matches_dgemm_dev_on_two_calls_second_smaller():
    rng = np.random.default_rng(4242)
    stream = native.stream()
    # (layout_a, layout_b, m, n, k, alpha, beta); the first two use split-K partials in the scratch
    # (few output tiles, long K), the second of each pair is smaller than the first.
    shapes = [(native.ROW, native.ROW, 256, 192, 50000, 1.25, 0.5),
              (native.COL, native.ROW, 128, 100, 20000, -0.75, 0.0),
              (native.ROW, native.COL, 640, 520, 700, 1.0, -1.5),
              (native.COL, native.COL, 77, 65, 33, 2.0, 1.0)]
    for la, lb, m, n, k, alpha, beta in shapes:
        A, B, C0, lda, ldb = _case(rng, la, lb, m, n, k)
        da, db = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(B)
        want = native.DeviceArray.from_numpy(C0)
        got = native.DeviceArray.from_numpy(C0)
        st = native.dgemm_dev(la, lb, m, n, k, alpha, da.ptr, lda, db.ptr, ldb, beta, want.ptr, n,
                              native.MODE_FP64, stream, stats=True)
        assert st.kernel_name == "mfma_f64"
        native.check(_noalloc(la, lb, m, n, k, alpha, da.ptr, lda, db.ptr, ldb, beta, got.ptr, n, stream))
        native.stream_synchronize(stream)
        w, g = want.to_numpy(), got.to_numpy()
        assert np.array_equal(w, g), (la, lb, m, n, k, float(np.max(np.abs(w - g))))
        # and against the plain product, at the reference's CPU tolerance scale
        opA = A.reshape(m, k) if la == native.ROW else A.reshape(k, m).T
        opB = B.reshape(k, n) if lb == native.ROW else B.reshape(n, k).T
        ref = alpha * (opA @ opB) + beta * C0.reshape(m, n)
        assert np.max(np.abs(g.reshape(m, n) - ref)) <= 1e-12 * k * max(1.0, np.max(np.abs(ref)))


def test_noalloc_null_stream_and_degenerate_shapes():
    rng = np.random.default_rng(7)
    m, n, k = 96, 80, 4096
    A, B, C0, lda, ldb = _case(rng, native.ROW, native.ROW, m, n, k)
    da, db = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(B)
    want, got = native.DeviceArray.from_numpy(C0), native.DeviceArray.from_numpy(C0)
    native.dgemm_dev(native.ROW, native.ROW, m, n, k, 0.5, da.ptr, lda, db.ptr, ldb, 2.0, want.ptr, n)
    native.check(_noalloc(native.ROW, native.ROW, m, n, k, 0.5, da.ptr, lda, db.ptr, ldb, 2.0, got.ptr, n, None))
    native.synchronize()
    assert np.array_equal(want.to_numpy(), got.to_numpy())
    # alpha == 0 scales C; m == 0 does nothing; bad arguments are rejected like jlc_dgemm_dev's
    native.check(_noalloc(native.ROW, native.ROW, m, n, k, 0.0, None, lda, None, ldb, 0.5, got.ptr, n, None))
    native.synchronize()
    assert np.array_equal(got.to_numpy(), 0.5 * want.to_numpy())
    native.check(_noalloc(native.ROW, native.ROW, 0, n, k, 1.0, None, lda, None, ldb, 0.0, None, n, None))
    assert _noalloc(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, lda, db.ptr, ldb, 0.0, got.ptr, n - 1,
                    None) == native.JLC_ERR_INVALID_ARG


def test_noalloc_rejects_partials_beyond_the_scratch():
    """With a tuning that splits K into many slices over many tiles (the round model switched off, so the
    full count is taken), the partials of a 2048^2 output (64 slices x 2048^2 doubles = 2 GiB) exceed the
    256 MiB scratch: the call fails with a message and launches nothing; jlc_dgemm_dev serves the same
    call from its own workspace."""
    old = native.tune_get("splitk_tiles_per_cu")
    old_quant = native.tune_get("splitk_quant")
    try:
        native.tune_set("splitk_tiles_per_cu", 64)
        native.tune_set("splitk_quant", 0)
        m = n = 2048
        k = 65536
        da, db, dc = native.DeviceArray(m * k), native.DeviceArray(k * n), native.DeviceArray(m * n)
        rc = _noalloc(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, dc.ptr, n, None)
        assert rc == native.JLC_ERR_UNSUPPORTED
        assert "scratch" in native.last_error()
    finally:
        native.tune_set("splitk_tiles_per_cu", old)
        native.tune_set("splitk_quant", old_quant)


def test_noalloc_skips_the_quantization_split_its_scratch_cannot_hold():
    """6000^3: jlc_dgemm_dev fills the grid's last round by splitting K three ways (partials 864 MB);
    the no-alloc call, whose scratch is 256 MiB, runs the same product unsplit instead of failing.
    The two results are the same product summed in another order: within the FP64 bound of each other,
    and the unsplit one equal bit for bit to jlc_dgemm_dev with the split switched off."""
    rng = np.random.default_rng(6000)
    m = n = k = 6000
    A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    da, db = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(B)
    split, got, plain = native.DeviceArray(m * n), native.DeviceArray(m * n), native.DeviceArray(m * n)
    native.dgemm_dev(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, split.ptr, n)
    native.check(_noalloc(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, got.ptr, n, None))
    native.synchronize()
    old = native.tune_get("splitk_quant")
    try:
        native.tune_set("splitk_quant", 0)
        native.dgemm_dev(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, plain.ptr, n)
    finally:
        native.tune_set("splitk_quant", old)
    g, p, s = got.to_numpy(), plain.to_numpy(), split.to_numpy()
    assert np.array_equal(g, p)
    assert not np.array_equal(g, s)  # the default call did split
    bound = 2 * k * 2.0 ** -53 * np.max(np.abs(A)) * np.max(np.abs(B)) * k
    assert np.max(np.abs(g - s)) <= bound
