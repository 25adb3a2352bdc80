"""Every measured configuration at its full size, on the GPU, through the C ABI and the host mirror.

Each test runs one configuration at the size it is measured at and checks it against the oracle or
against what the mathematics fixes, at sizes where a full CPU reference would take minutes:

* config 1: 512^3, alpha = 1.5, beta = 0.5, entries uniform[-1,1], through OptimizedBLAS3.gemm and
  BLAS3Kernels.gemm under the default (tuned) policy, against GemmReference at the reference's CPU
  tolerance (GemmParityTest.java:38-72; GemmReference.java:175-177).
* configs 2a-2c: 8192^3 row-major; gemmStridedTransA 4096 x 8192 x 16384 with A stored 16384 x 4096;
  gemmStridedColMajorA / ColMajorB at 6000^3 (not a multiple of the tile). Small-integer inputs make
  every product exact in any summation order, so 64 full rows and 64 full columns of C are compared
  bit for bit with numpy, and the row and column checksums of the whole C (C 1 = A (B 1),
  1^T C = (1^T A) B) exactly. Then uniform[-1,1] inputs: 256 sampled entries against double-double
  dot products, inside the componentwise bound k u (|A||B|)_ij, relative error of the sample <= 1e-12.
* config 2d: matvec 32768 x 32768 with small-integer inputs: y bit for bit, through jlc_dgemv_dev and
  through jlc_dgemm_dev's n == 1 regime.
* config 3: 16384^3, a = (u - 0.5) exp(phi g) for phi in {0.5, 2, 8}: each automatic split form runs
  its own kernels with a max relative error on the 64 x 64 double-double sample at most the FP64
  kernel's, or runs the FP64 kernel (the form's picks are pinned per phi).
* config 4: 100,000 uneven problems (M, N, K log-uniform on [2, 256], 10% 2/3/4 cubes, alpha and beta
  uniform on [-2, 2]) in one grouped launch, FP64 and the INT8 slice plan: every 10th problem against
  the oracle (GemmReference at its CPU tolerance; the slice restatement bit for bit) and every
  problem's checksum inside the summed componentwise bound.
* config 5a: blocked QR of a 32768 x 4096 A = U diag(s) V^T (s geometric 1 -> 1e-8), FP64 panels with
  FP64, CRT and INT8-slice trailing updates: ||A - QR||/||A|| and ||Q^T Q - I|| inside the reference's
  size-graded tolerance (DecompositionTests.java:183-188) and inside FP64-level bounds.
* config 5b: 20,000 QRs (m uniform on [16, 192], n = m/2, same construction) in one graph replay:
  every matrix's tau and R against the HouseholderQR restatement, and its reconstruction and
  orthogonality errors.
"""
from __future__ import annotations

import sys
import time
import zlib
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np
import pytest

from jlc_amd import native
from jlc_amd.compute import BLAS3Kernels, DispatchPolicy, Matrix, OptimizedBLAS3
from jlc_amd.workloads import BatchedWorkload, batched_shapes
from oracle import oracle as orc

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

pytestmark = pytest.mark.gpu

U = 2.0 ** -53
ROW, COL = native.ROW, native.COL
POOL = 8


def log(msg: str) -> None:
    print(f"[configs] {msg}", flush=True)


def small_ints(rng, count: int) -> np.ndarray:
    """Integers in [-8, 8] as float64: products of such matrices are exact in any order below 2^53."""
    return rng.integers(-8, 9, size=count, dtype=np.int8).astype(np.float64)


def op(x: np.ndarray, layout: int, rows: int, cols: int) -> np.ndarray:
    """The rows x cols operand an array stored in `layout` denotes (a view)."""
    return x.reshape(rows, cols) if layout == ROW else x.reshape(cols, rows).T


def fetch_rows(dc: native.DeviceArray, n: int, rows) -> np.ndarray:
    out = np.empty((len(rows), n))
    lib = native.load()
    for t, r in enumerate(rows):
        native.check(lib.jlc_memcpy_d2h(out[t].ctypes.data, dc.ptr + 8 * int(r) * n, 8 * n))
    return out


# ---------------------------------------------------------------------------------------------
# config 1

def test_config1_512_alpha_beta_through_the_api():
    m = n = k = 512
    alpha, beta = 1.5, 0.5
    a = Matrix.wrap(orc.uniform_array(m * k, 20260611), m, k)
    b = Matrix.wrap(orc.uniform_array(k * n, 20260612), k, n)
    c0 = orc.uniform_array(m * n, 20260613)
    expected = orc.gemm(a.getRawData(), b.getRawData(), c0, m, k, n, alpha, beta)
    policy = DispatchPolicy.defaultPolicy()  # the tuned gate sends 2.7e8 flops to HIP
    for name, fn in (("OptimizedBLAS3.gemm", OptimizedBLAS3.gemm), ("BLAS3Kernels.gemm", BLAS3Kernels.gemm)):
        c = Matrix.wrap(c0.copy(), m, n)
        fn(a, b, c, alpha, beta, policy)
        orc.assert_parity(name, m, k, n, alpha, beta, 20260611, expected, c.getRawData(), 1e-12,
                          orc.cpu_abs_tolerance(expected, k))


# ---------------------------------------------------------------------------------------------
# configs 2a-2c

# (name, layout_a, layout_b, m, n, k, host-mirror call)
LAYOUT_CASES = [
    ("2a_rowmajor_8192", ROW, ROW, 8192, 8192, 8192, "gemmStrided"),
    ("2b_transA_4096x8192x16384", COL, ROW, 4096, 8192, 16384, "gemmStridedTransA"),
    ("2c_colmajorA_6000", COL, ROW, 6000, 6000, 6000, "gemmStridedColMajorA"),
    ("2c_colmajorB_6000", ROW, COL, 6000, 6000, 6000, "gemmStridedColMajorB"),
]


def _mirror_call(call: str, a, lda, b, ldb, c, ldc, m, n, k):
    """The reference's strided call for this layout (BLAS3Kernels.java:244-303): note the argument
    order m, k, n, and that gemmStridedTransA's first size is the inner dimension."""
    fn = getattr(BLAS3Kernels, call)
    if call == "gemmStridedTransA":
        fn(a, 0, lda, b, 0, ldb, c, 0, ldc, k, m, n, 1.0, 0.0)
    else:
        fn(a, 0, lda, b, 0, ldb, c, 0, ldc, m, k, n, 1.0, 0.0)


@pytest.mark.parametrize("name,la,lb,m,n,k,call", LAYOUT_CASES, ids=[c[0] for c in LAYOUT_CASES])
def test_config2_layouts_exact_integers_and_sampled_dd(name, la, lb, m, n, k, call):
    t0 = time.perf_counter()
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    lda = k if la == ROW else m
    ldb = n if lb == ROW else k
    # exact: small integers through the host mirror (the host-pointer entry point)
    A = small_ints(rng, m * k)
    B = small_ints(rng, k * n)
    C = np.full(m * n, np.nan)
    _mirror_call(call, A, lda, B, ldb, C, n, m, n, k)
    opA, opB, Cm = op(A, la, m, k), op(B, lb, k, n), C.reshape(m, n)
    rows = np.sort(rng.choice(m, 64, replace=False))
    cols = np.sort(rng.choice(n, 64, replace=False))
    assert np.array_equal(Cm[rows], opA[rows] @ opB), name
    assert np.array_equal(Cm[:, cols], opA @ np.ascontiguousarray(opB[:, cols])), name
    assert np.array_equal(Cm.sum(axis=1), opA @ opB.sum(axis=1)), name
    assert np.array_equal(Cm.sum(axis=0), opA.sum(axis=0) @ opB), name
    del A, B, C, opA, opB, Cm
    # uniform[-1,1] on device pointers: sampled entries against double-double dot products
    A = rng.uniform(-1.0, 1.0, m * k)
    B = rng.uniform(-1.0, 1.0, k * n)
    da, db, dc = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(B), native.DeviceArray(m * n)
    st = native.dgemm_dev(la, lb, m, n, k, 1.0, da.ptr, lda, db.ptr, ldb, 0.0, dc.ptr, n, native.MODE_FP64)
    assert st.kernel_name == "mfma_f64"
    ii = rng.integers(0, m, 256)
    jj = rng.integers(0, n, 256)
    hi, lo, absum = orc.dd_entries(la, lb, k, A, lda, B, ldb, ii, jj)
    got = dc.to_numpy().reshape(m, n)[ii, jj]
    err = np.abs(got - hi - lo)
    assert np.all(err <= k * U * absum), (name, float(np.max(err / (U * absum))))
    assert np.linalg.norm(err) <= 1e-12 * np.linalg.norm(hi), name
    log(f"{name}: exact and sampled checks in {time.perf_counter() - t0:.1f}s, "
        f"max componentwise err {np.max(err / (U * absum)):.2f} u")


# ---------------------------------------------------------------------------------------------
# config 2d

def test_config2d_matvec_32768_exact_integers():
    t0 = time.perf_counter()
    mv = 32768
    rng = np.random.default_rng(2032768)
    x = small_ints(rng, mv)
    dA = native.DeviceArray(mv * mv)
    lib = native.load()
    y_exp = np.empty(mv)
    chunk = 2048
    for r0 in range(0, mv, chunk):
        blk = small_ints(rng, chunk * mv)
        native.check(lib.jlc_memcpy_h2d(dA.ptr + 8 * r0 * mv, blk.ctypes.data, blk.size * 8))
        y_exp[r0:r0 + chunk] = blk.reshape(chunk, mv) @ x
    dx, dy = native.DeviceArray.from_numpy(x), native.DeviceArray(mv)
    st = native.dgemv_dev(mv, mv, 1.0, dA.ptr, mv, dx.ptr, 1, 0.0, dy.ptr, 1)
    assert st.kernel_name == "gemv"
    assert np.array_equal(dy.to_numpy(), y_exp)
    # the same product as an n == 1 GEMM (the MATVEC regime of the dispatch)
    dy2 = native.DeviceArray.from_numpy(np.full(mv, np.nan))
    st = native.dgemm_dev(ROW, ROW, mv, 1, mv, 1.0, dA.ptr, mv, dx.ptr, 1, 0.0, dy2.ptr, 1)
    assert st.kernel_name == "gemv"
    assert np.array_equal(dy2.to_numpy(), y_exp)
    log(f"2d matvec 32768^2: bitwise in {time.perf_counter() - t0:.1f}s")


# ---------------------------------------------------------------------------------------------
# config 3

# Automatic picks per phi on the 16384^3 workload: (kernel, slices or moduli) for each split form;
# ("mfma_f64", 0) is the FP64 kernel serving a spread the form cannot hold.
CONFIG3_PICKS = {
    0.5: {"split_crt": ("split_crt", 15), "split_int8": ("split_int8", 9), "split_bf16": ("split_bf16", 8)},
    2.0: {"split_crt": ("split_crt", 16), "split_int8": ("mfma_f64", 0), "split_bf16": ("mfma_f64", 0)},
    8.0: {"split_crt": ("mfma_f64", 0), "split_int8": ("mfma_f64", 0), "split_bf16": ("mfma_f64", 0)},
}


@pytest.mark.parametrize("phi", [0.5, 2.0, 8.0])
def test_config3_16384_split_forms_against_double_double(phi):
    import bench
    t0 = time.perf_counter()
    m = n = k = 16384
    seed = 20260603
    rng = np.random.default_rng(seed)
    a = bench.make_operand(rng, m, k, phi)
    b = bench.make_operand(rng, k, n, phi)
    srng = np.random.default_rng(seed + 1)
    rows = np.sort(srng.choice(m, 64, replace=False))
    cols = np.sort(srng.choice(n, 64, replace=False))
    As = a.reshape(m, k)[rows].copy().reshape(-1)
    Bs = np.ascontiguousarray(b.reshape(k, n)[:, cols]).reshape(-1)
    ii, jj = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
    hi, lo, absum = orc.dd_entries(ROW, ROW, k, As, k, Bs, 64, ii.ravel(), jj.ravel())
    exact = hi + lo
    nz = exact != 0.0
    da, db = native.DeviceArray.from_numpy(a), native.DeviceArray.from_numpy(b)
    del a, b
    dc = native.DeviceArray(m * n)
    log(f"config 3 phi={phi}: inputs and dd sample in {time.perf_counter() - t0:.1f}s")

    def run(mode):
        st = native.dgemm_dev(ROW, ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, dc.ptr, n, mode)
        got = fetch_rows(dc, n, rows)[:, cols].reshape(-1)
        err = np.abs(got - hi - lo)
        return st, float(np.max(err[nz] / np.abs(exact[nz]))), float(np.max(err / absum) / U)

    _, rel64, comp64 = run(native.MODE_FP64)
    picks = {}
    for form, mode in (("split_crt", native.MODE_SPLIT_CRT), ("split_int8", native.MODE_SPLIT_INT8),
                       ("split_bf16", native.MODE_SPLIT_BF16)):
        st, rel, comp = run(mode)
        picks[form] = (st.kernel_name, st.slices)
        log(f"config 3 phi={phi}: {form} -> {st.kernel_name} ({st.slices}) max rel {rel:.3e} "
            f"(FP64 {rel64:.3e}), componentwise {comp:.2f} u (FP64 {comp64:.2f} u)")
        if st.kernel_name == "mfma_f64":
            assert st.slices == 0
            assert rel == rel64  # the same kernel on the same inputs
        else:
            assert st.kernel_name == form
            assert rel <= rel64, (phi, form, rel, rel64)
    assert picks == CONFIG3_PICKS[phi], picks


# ---------------------------------------------------------------------------------------------
# config 4

def _checksum_bound(A, B, C0, Cg, m, n, k, alpha, beta, extra=0.0):
    """|sum(C_gpu) - sum(alpha A B + beta C0)| may not exceed the componentwise bound summed over C,
    (k + 2) u sum_ij (|alpha| |A||B| + |beta| |C0|)_ij, plus the rounding of the three sums."""
    absA, absB = np.abs(A.reshape(m, k)), np.abs(B.reshape(k, n))
    s_abs = abs(alpha) * float(absA.sum(axis=0) @ absB.sum(axis=1)) + abs(beta) * float(np.abs(C0).sum())
    ref = alpha * float(A.reshape(m, k).sum(axis=0) @ B.reshape(k, n).sum(axis=1)) + beta * float(C0.sum())
    tol = (k + 2 + 4 * (m * n + k)) * U * (s_abs + float(np.abs(Cg).sum())) + extra
    return abs(float(Cg.sum()) - ref), tol


def test_config4_100k_batched_fp64_and_int8_slices():
    t0 = time.perf_counter()
    dims, alpha, beta = batched_shapes(100000, 20260603)
    w = BatchedWorkload(dims, alpha, beta, seed=20260604)
    descs = w.descs()
    log(f"config 4: {w.count} problems set up in {time.perf_counter() - t0:.1f}s")
    S = 8
    results = {}
    for name, mode in (("fp64", native.MODE_FP64), ("split_int8", native.mode_slices(native.MODE_SPLIT_INT8, S))):
        w.reset_c()
        native.synchronize()
        plan = native.GroupedPlan(descs, mode)
        plan.replay()
        native.synchronize()
        results[name] = w.dc.to_numpy()
        plan.destroy()
    c64, ci8 = results["fp64"], results["split_int8"]

    def check_every_10th(i):
        m, n, k = (int(x) for x in w.dims[i])
        A, B, C0, Cg = w.problem(i, c64)
        al, be = float(w.alpha[i]), float(w.beta[i])
        exp = C0.copy()
        orc.gemm_general(ROW, ROW, m, n, k, al, A, k, B, n, be, exp, n, nthreads=1)
        orc.assert_parity(f"grouped fp64 problem {i}", m, k, n, al, be, i, exp, Cg, 1e-12,
                          orc.cpu_abs_tolerance(exp, k))
        exp8 = C0.copy()
        orc.split_gemm(ROW, ROW, m, n, k, al, A, k, B, n, be, exp8, n, S, nthreads=1)
        assert np.array_equal(ci8[int(w.c_off[i]):int(w.c_off[i]) + m * n], exp8), i

    with ThreadPoolExecutor(POOL) as ex:
        list(ex.map(check_every_10th, range(0, w.count, 10)))
    log(f"config 4: every 10th problem against the oracle, {time.perf_counter() - t0:.1f}s")
    worst = {"fp64": 0.0, "split_int8": 0.0}
    for i in range(w.count):
        m, n, k = (int(x) for x in w.dims[i])
        A, B, C0, _ = w.problem(i)
        al, be = float(w.alpha[i]), float(w.beta[i])
        lo_, hi_ = int(w.c_off[i]), int(w.c_off[i]) + m * n
        d, tol = _checksum_bound(A, B, C0, c64[lo_:hi_], m, n, k, al, be)
        assert d <= tol, ("fp64", i, d, tol)
        worst["fp64"] = max(worst["fp64"], d / tol)
        # slices: each entry truncated below 2^(e - 7S) of its row / column scale 2^e (e the frexp
        # exponent of the row / column max), and the dropped digit pairs below that again
        amax = np.abs(A.reshape(m, k)).max(axis=1)
        bmax = np.abs(B.reshape(k, n)).max(axis=0)
        trunc = abs(al) * S * S * k * 4.0 * float(amax.sum() * bmax.sum()) * 2.0 ** (-7 * S)
        d, tol = _checksum_bound(A, B, C0, ci8[lo_:hi_], m, n, k, al, be, trunc)
        assert d <= tol, ("split_int8", i, d, tol)
        worst["split_int8"] = max(worst["split_int8"], d / tol)
    log(f"config 4: every checksum inside its bound (worst fraction {worst}), {time.perf_counter() - t0:.1f}s")


# ---------------------------------------------------------------------------------------------
# configs 5a, 5b

def _reference_tolerance(n: int) -> float:
    """DecompositionTests.getTolerance (DecompositionTests.java:183-188)."""
    if n <= 10:
        return 1e-10 * np.sqrt(n)
    if n <= 30:
        return 1e-8 * np.sqrt(n)
    if n <= 100:
        return 1e-6 * np.sqrt(n)
    return 1e-5 * np.sqrt(n)


def test_config5a_tall_qr_32768x4096_all_update_forms():
    t0 = time.perf_counter()
    m, n, nb = 32768, 4096, 128
    rng = np.random.default_rng(5032768)

    def q_of_gaussian(r, c):
        g = native.DeviceArray.from_numpy(rng.standard_normal(r * c))  # column-major r x c
        tau = native.DeviceArray(c)
        native.dgeqrf_dev(r, c, g.ptr, r, tau.ptr, nb)
        q = native.DeviceArray(r * c)
        native.dorgqr_dev(r, c, g.ptr, r, tau.ptr, nb, q.ptr, r)
        return q

    Ud = q_of_gaussian(m, n)
    V = q_of_gaussian(n, n).to_numpy().reshape(n, n)          # row j = column j of V
    Vs = (V.T * np.geomspace(1.0, 1e-8, n)).T.copy()
    dVs = native.DeviceArray.from_numpy(Vs.reshape(-1))
    A = native.DeviceArray(m * n)
    native.dgemm_dev(COL, ROW, n, m, n, 1.0, dVs.ptr, n, Ud.ptr, m, 0.0, A.ptr, m)  # A^T = (V diag(s)) U^T
    work, tau = native.DeviceArray(m * n), native.DeviceArray(n)
    tol = _reference_tolerance(n)
    for name, mode in (("fp64", native.MODE_FP64), ("split_crt", native.MODE_SPLIT_CRT),
                       ("split_int8", native.MODE_SPLIT_INT8)):
        native.dev_copy(work.ptr, A.ptr, m * n * 8)
        st = native.dgeqrf_dev(m, n, work.ptr, m, tau.ptr, nb, mode)
        assert st.kernel_name == "qr"
        recon, orth = native.dqr_errors_dev(m, n, A.ptr, m, work.ptr, m, tau.ptr, nb)
        log(f"config 5a {name}: recon {recon:.3e} orth {orth:.3e} ({st.kernel_ms:.1f} ms)")
        assert recon < tol and orth < tol, (name, recon, orth, tol)
        assert recon < 1e-13 and orth < 1e-11, (name, recon, orth)
    log(f"config 5a: {time.perf_counter() - t0:.1f}s")


def test_config5b_20000_batched_qrs_every_matrix_against_the_oracle():
    t0 = time.perf_counter()
    rng = np.random.default_rng(5020000)
    count = 20000
    ms_ = rng.integers(16, 193, count)
    mats = [None] * count
    for mm in np.unique(ms_):
        idx = np.nonzero(ms_ == mm)[0]
        nn = max(1, int(mm) // 2)
        Uq, _ = np.linalg.qr(rng.standard_normal((len(idx), int(mm), nn)))
        Vq, _ = np.linalg.qr(rng.standard_normal((len(idx), nn, nn)))
        Ab = (Uq * np.geomspace(1.0, 1e-8, nn)) @ np.swapaxes(Vq, 1, 2)
        for t, i in enumerate(idx):
            mats[i] = Ab[t]
    flat = np.concatenate([np.asfortranarray(M).reshape(-1, order="F") for M in mats])
    offs = np.concatenate([[0], np.cumsum([M.size for M in mats])[:-1]])
    toffs = np.concatenate([[0], np.cumsum([min(M.shape) for M in mats])[:-1]])
    da = native.DeviceArray.from_numpy(flat)
    dt = native.DeviceArray(int(toffs[-1] + min(mats[-1].shape)))
    descs = [native.QrDesc(M.shape[0], M.shape[1], M.shape[0], da.at(int(o)), dt.at(int(t)))
             for M, o, t in zip(mats, offs, toffs)]
    plan = native.QrBatchedPlan(descs)
    plan.replay()
    native.synchronize()
    got, taus = da.to_numpy(), dt.to_numpy()
    plan.destroy()
    log(f"config 5b: {count} QRs factored, {time.perf_counter() - t0:.1f}s")

    def check(i):
        M = mats[i]
        m, n = M.shape
        f = got[offs[i]:offs[i] + m * n].copy()
        tau = taus[toffs[i]:toffs[i] + n]
        ref = np.asfortranarray(M).reshape(-1, order="F").copy()
        tau_ref = orc.qr_factor(ref, m, n)
        R, Rr = np.triu(f.reshape(n, m).T[:n]), np.triu(ref.reshape(n, m).T[:n])
        # tau_j is fixed by column j of the partly reduced matrix, whose norm is |R_jj|: rounding of
        # size u ||M|| in that column moves tau_j by about u ||M|| / |R_jj| relative, up to ~1e-8 here
        # where the singular values reach 1e-8, whatever the (backward-stable) summation order.
        rtol = 1e-10 + 64.0 * U * np.linalg.norm(M) / np.maximum(np.abs(np.diag(Rr)), 1e-300)
        assert np.all(np.abs(tau - tau_ref) <= rtol * np.abs(tau_ref) + 1e-12), i
        np.testing.assert_allclose(R, Rr, rtol=0, atol=1e-10 * np.max(np.abs(Rr)))
        Q = orc.qr_form_q(f, tau, m, n).reshape(n, m).T
        return (float(np.linalg.norm(M - Q @ R) / np.linalg.norm(M)), float(np.linalg.norm(Q.T @ Q - np.eye(n))))

    with ThreadPoolExecutor(POOL) as ex:
        errs = np.array(list(ex.map(check, range(count))))
    log(f"config 5b: max recon {errs[:, 0].max():.3e}, max orth {errs[:, 1].max():.3e}, "
        f"{time.perf_counter() - t0:.1f}s")
    assert errs[:, 0].max() < 1e-13 and errs[:, 1].max() < 1e-13
