"""GPU parity of the grouped (batched, uneven) FP64 GEMM against the oracle."""
import numpy as np
import pytest

from jlc_amd import native
from jlc_amd.workloads import BatchedWorkload, batched_shapes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _check_problem(w, i, c_now, tag):
    m, n, k = (int(x) for x in w.dims[i])
    la, lb = (int(x) for x in w.layouts[i])
    lda, ldb = w.lds(i)
    A, B, C0, Cn = w.problem(i, c_now)
    exp = C0.copy()
    orc.gemm_general(la, lb, m, n, k, float(w.alpha[i]), A, lda, B, ldb, float(w.beta[i]), exp, n)
    orc.assert_parity(f"{tag} problem {i}", m, k, n, float(w.alpha[i]), float(w.beta[i]), i, exp, Cn, 1e-12,
                      orc.cpu_abs_tolerance(exp, k))


def test_grouped_uneven_all_layouts_against_oracle():
    dims, alpha, beta = batched_shapes(1500, seed=41, lo=1, hi=300)
    rng = np.random.default_rng(42)
    layouts = rng.integers(0, 2, size=(len(dims), 2))
    w = BatchedWorkload(dims, alpha, beta, seed=43, layouts=layouts)
    st = native.dgemm_grouped_dev(w.descs())
    assert st.kernel_name == "grouped"
    c = w.dc.to_numpy()
    for i in range(w.count):
        _check_problem(w, i, c, "grouped")


def test_grouped_degenerate_problems():
    dims = np.array([[0, 5, 3], [4, 0, 3], [5, 4, 0], [3, 3, 3], [1, 1, 1], [7, 2, 9]])
    alpha = np.array([1.0, 1.0, 2.5, 0.0, 1.0, -1.0])
    beta = np.array([0.0, 0.5, -0.5, 2.5, 0.0, 0.0])
    w = BatchedWorkload(dims, alpha, beta, seed=5)
    # alpha == 0 must not read A or B: poison problem 3's operands with NaN.
    A3 = w.a_off[3]
    w.a_host[A3:A3 + 9] = np.nan
    w.da = native.DeviceArray.from_numpy(w.a_host)
    # beta == 0 must not read C: poison problem 4's and 5's C.
    w.c_host[w.c_off[4]] = np.nan
    w.c_host[w.c_off[5]:w.c_off[5] + 14] = np.nan
    w.dc = native.DeviceArray.from_numpy(w.c_host)
    native.dgemm_grouped_dev(w.descs())
    c = w.dc.to_numpy()
    _, _, C0, C2 = w.problem(2, c)
    assert np.array_equal(C2, -0.5 * C0)            # k == 0: C = beta*C exactly
    _, _, C0, C3 = w.problem(3, c)
    assert np.array_equal(C3, 2.5 * C0)             # alpha == 0: A unread
    for i in (4, 5):
        m, n, k = (int(x) for x in w.dims[i])
        A, B, _, Ci = w.problem(i, c)
        assert np.all(np.isfinite(Ci))
        exp = np.zeros(m * n)  # the oracle multiplies C by beta even when beta == 0
        orc.gemm_general(0, 0, m, n, k, float(alpha[i]), A, k, B, n, 0.0, exp, n)
        orc.assert_parity(f"degenerate {i}", m, k, n, float(alpha[i]), 0.0, i, exp, Ci, 1e-12,
                          orc.cpu_abs_tolerance(exp, k))


def test_grouped_split_degenerate_problems():
    # The INT8 slice form keeps the same BLAS semantics: alpha == 0 or k == 0 scale C without reading
    # A or B, beta == 0 never reads C (NaN planted there stays out of the result).
    dims = np.array([[0, 5, 3], [4, 0, 3], [5, 4, 0], [3, 3, 3], [1, 1, 1], [7, 2, 9], [40, 33, 70]])
    alpha = np.array([1.0, 1.0, 2.5, 0.0, 1.0, -1.0, 1.5])
    beta = np.array([0.0, 0.5, -0.5, 2.5, 0.0, 0.0, 0.0])
    w = BatchedWorkload(dims, alpha, beta, seed=6)
    A3 = w.a_off[3]
    w.a_host[A3:A3 + 9] = np.nan
    w.da.copy_from(w.a_host)
    w.c_host[w.c_off[4]] = np.nan
    w.c_host[w.c_off[5]:w.c_off[5] + 14] = np.nan
    w.c_host[w.c_off[6]:w.c_off[6] + 40 * 33] = np.inf
    w.dc.copy_from(w.c_host)
    native.dgemm_grouped_dev(w.descs(), native.mode_slices(native.MODE_SPLIT_INT8, 7))
    c = w.dc.to_numpy()
    _, _, C0, C2 = w.problem(2, c)
    assert np.array_equal(C2, -0.5 * C0)
    _, _, C0, C3 = w.problem(3, c)
    assert np.array_equal(C3, 2.5 * C0)
    for i in (4, 5, 6):
        m, n, k = (int(x) for x in w.dims[i])
        A, B, _, Ci = w.problem(i, c)
        exp = np.zeros(m * n)
        orc.split_gemm(0, 0, m, n, k, float(alpha[i]), A, k, B, n, 0.0, exp, n, 7)
        assert np.all(np.isfinite(Ci)) and np.array_equal(Ci, exp), i


def test_graph_replays_match_eager_run():
    dims, alpha, beta = batched_shapes(800, seed=7)
    beta[:] = 0.0  # replays are then idempotent: every replay must give the same bits
    w = BatchedWorkload(dims, alpha, beta, seed=8)
    plan = native.GroupedPlan(w.descs())
    plan.run()
    native.synchronize()
    eager = w.dc.to_numpy()
    for rep in range(4):
        w.reset_c()                 # eager work between replays
        native.synchronize()
        plan.replay()
        native.synchronize()
        assert np.array_equal(w.dc.to_numpy(), eager), rep
    plan.destroy()
    for i in range(0, w.count, 37):
        _check_problem(w, i, eager, "graph")


@pytest.mark.parametrize("xcds,wps", [(1, 0), (8, 0), (8, 1), (3, 0)])
def test_tile_placement_does_not_change_results(xcds, wps):
    # The tile list's placement (per-XCD dealing, skip entries, grid width)
    # only moves tiles between wavefronts: every tile is computed once, by the
    # same instructions, so the bits must match the default placement's.
    dims, alpha, beta = batched_shapes(600, seed=11, lo=1, hi=300)
    ref_w = BatchedWorkload(dims, alpha, beta, seed=12)
    native.dgemm_grouped_dev(ref_w.descs())
    ref = ref_w.dc.to_numpy()
    old = (native.tune_get("grouped_xcds"), native.tune_get("grouped_waves_per_simd"))
    try:
        native.tune_set("grouped_xcds", xcds)
        native.tune_set("grouped_waves_per_simd", wps)
        w = BatchedWorkload(dims, alpha, beta, seed=12)
        native.dgemm_grouped_dev(w.descs())
        assert np.array_equal(w.dc.to_numpy(), ref)
    finally:
        native.tune_set("grouped_xcds", old[0])
        native.tune_set("grouped_waves_per_simd", old[1])
    for i in range(0, w.count, 29):
        _check_problem(w, i, ref, "placement")


@pytest.mark.parametrize("small", [0, 8, 16, 300])
def test_tile_shapes_do_not_change_results(small):
    # Problems with m (n) up to grouped_small_tile run 16-row (16-column) tiles; every 16 x 16 block of C
    # is the same MFMA chain over k in either tile shape, so the bits match the all-32 x 32 run, and the
    # mixed-shape lists (0 / 16-only / all shapes) are checked against the oracle too.
    dims, alpha, beta = batched_shapes(900, seed=19, lo=1, hi=70)
    rng = np.random.default_rng(20)
    layouts = rng.integers(0, 2, size=(len(dims), 2))
    old = native.tune_get("grouped_small_tile")
    try:
        native.tune_set("grouped_small_tile", 0)
        ref_w = BatchedWorkload(dims, alpha, beta, seed=21, layouts=layouts)
        native.dgemm_grouped_dev(ref_w.descs())
        ref = ref_w.dc.to_numpy()
        native.tune_set("grouped_small_tile", small)
        w = BatchedWorkload(dims, alpha, beta, seed=21, layouts=layouts)
        native.dgemm_grouped_dev(w.descs())
        assert np.array_equal(w.dc.to_numpy(), ref)
    finally:
        native.tune_set("grouped_small_tile", old)
    for i in range(0, w.count, 7):
        _check_problem(w, i, ref, f"small tiles {small}")


def _check_split_problem(w, i, c_now, slices, tag):
    m, n, k = (int(x) for x in w.dims[i])
    la, lb = (int(x) for x in w.layouts[i])
    lda, ldb = w.lds(i)
    A, B, C0, Cn = w.problem(i, c_now)
    exp = C0.copy()
    alpha, beta = float(w.alpha[i]), float(w.beta[i])
    if alpha == 0.0 or k == 0:
        exp = np.zeros_like(C0) if beta == 0.0 else beta * C0   # C = beta * C, op(A), op(B) unread
    elif m and n:
        orc.split_gemm(la, lb, m, n, k, alpha, A, lda, B, ldb, beta, exp, n, slices)
    same = (Cn == exp) | (np.isnan(Cn) & np.isnan(exp))
    assert same.all(), (tag, i, m, n, k, la, lb, int((~same).sum()))


@pytest.mark.parametrize("slices", [3, 5, 7, 8, 9])
def test_grouped_split_int8_bitwise_against_restatement(slices):
    # Each problem must carry the single-GEMM INT8 slice form's bits: same
    # exponents, digits, diagonal sums and FP64 recombination.
    dims, alpha, beta = batched_shapes(700, seed=60 + slices, lo=1, hi=300)
    rng = np.random.default_rng(61 + slices)
    layouts = rng.integers(0, 2, size=(len(dims), 2))
    alpha[::50] = 0.0                  # alpha == 0: C = beta * C
    w = BatchedWorkload(dims, alpha, beta, seed=62 + slices, layouts=layouts)
    # wide exponent ranges inside some problems, and Inf / NaN rows and columns in others
    for i in range(0, w.count, 17):
        m, n, k = (int(x) for x in dims[i])
        a0 = int(w.a_off[i])
        w.a_host[a0:a0 + m * k] *= np.exp2(rng.integers(-40, 40, m * k))
    for i in range(3, w.count, 41):
        m, n, k = (int(x) for x in dims[i])
        if m * k:
            w.a_host[int(w.a_off[i]) + rng.integers(0, m * k)] = np.inf
        if k * n:
            w.b_host[int(w.b_off[i]) + rng.integers(0, k * n)] = np.nan
    w.da.copy_from(w.a_host)
    w.db.copy_from(w.b_host)
    st = native.dgemm_grouped_dev(w.descs(), native.mode_slices(native.MODE_SPLIT_INT8, slices))
    assert st.kernel_name == "grouped" and st.slices == slices
    c = w.dc.to_numpy()
    for i in range(w.count):
        _check_split_problem(w, i, c, slices, f"split{slices}")


def test_grouped_split_int8_chunked_depth_and_replay():
    # K = 15000 crosses the INT32 chunk (462 blocks of 32 at 9 slices): one mid-K flush.
    dims = np.array([[40, 33, 15000], [5, 70, 14785], [64, 64, 300], [3, 3, 3]])
    alpha = np.array([1.5, -0.5, 2.0, 1.0])
    beta = np.array([0.5, 0.0, -1.0, 0.25])
    w = BatchedWorkload(dims, alpha, beta, seed=70)
    plan = native.GroupedPlan(w.descs(), native.mode_slices(native.MODE_SPLIT_INT8, 9))
    plan.run()
    native.synchronize()
    first = w.dc.to_numpy()
    for i in range(w.count):
        _check_split_problem(w, i, first, 9, "chunked")
    for rep in range(3):
        w.reset_c()
        native.synchronize()
        plan.replay()
        native.synchronize()
        assert np.array_equal(w.dc.to_numpy(), first), rep
    plan.destroy()


def test_grouped_split_int8_slice_count_checked():
    d = native.GemmDesc(0, 0, 4, 4, 4, 0, 1.0, 0.0, None, None, None, 4, 4, 4)
    for mode in (native.MODE_SPLIT_INT8, native.mode_slices(native.MODE_SPLIT_INT8, 2),
                 native.mode_slices(native.MODE_SPLIT_INT8, 10), native.MODE_SPLIT_CRT):
        with pytest.raises(native.JlcError):
            native.GroupedPlan([d], mode)


def test_invalid_descriptor_rejected():
    d = native.GemmDesc(0, 0, 4, 4, 4, 0, 1.0, 0.0, None, None, None, 4, 4, 4)
    with pytest.raises(native.JlcError) as ei:
        native.GroupedPlan([d])
    assert ei.value.status == native.JLC_ERR_INVALID_ARG and "descriptor 0" in ei.value.message
