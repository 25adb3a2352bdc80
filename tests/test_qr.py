"""GPU blocked Householder QR and batched small QR against the oracle and the
reference's QR acceptance tests (src/test/java/net/faulj/stress/HouseholderQRTest.java,
src/test/java/net/faulj/decomposition/DecompositionTests.java:183-188, :459-527)."""
import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _factor(A, nb=128, mode=native.MODE_FP64):
    m, n = A.shape
    a_cm = np.asfortranarray(A).reshape(-1, order="F").copy()
    da = native.DeviceArray.from_numpy(a_cm)
    d0 = native.DeviceArray.from_numpy(a_cm)
    dt = native.DeviceArray(min(m, n))
    st = native.dgeqrf_dev(m, n, da.ptr, m, dt.ptr, nb, mode)
    return da, d0, dt, st


def _errors(A, da, d0, dt, nb):
    m, n = A.shape
    return native.dqr_errors_dev(m, n, d0.ptr, m, da.ptr, m, dt.ptr, nb)


def test_reference_square_cases_n_1_to_100():
    """HouseholderQRTest: n = 1..100 random square, recon and ortho < 1e-12."""
    rng = np.random.default_rng(12)
    for n in list(range(1, 21)) + [33, 64, 65, 100]:
        A = rng.uniform(-1, 1, (n, n))
        da, d0, dt, _ = _factor(A, nb=32)
        recon, orth = _errors(A, da, d0, dt, 32)
        assert recon < 1e-12 and orth < 1e-12, (n, recon, orth)


def test_reference_thin_case_and_oracle_parity():
    """Thin 100 x 40 (HouseholderQRTest.java:100-121): same R and tau as the oracle restatement."""
    rng = np.random.default_rng(40)
    for (m, n, nb) in [(100, 40, 16), (300, 300, 64), (513, 200, 128), (1000, 640, 128)]:
        A = rng.standard_normal((m, n))
        da, d0, dt, st = _factor(A, nb=nb)
        assert st.kernel_name == "qr"
        recon, orth = _errors(A, da, d0, dt, nb)
        tol = 1e-10 * np.sqrt(n)  # size-graded (DecompositionTests.java:183-188)
        assert recon < tol and orth < tol, (m, n, recon, orth)
        ref = np.asfortranarray(A).reshape(-1, order="F").copy()
        tau_ref = orc.qr_factor(ref, m, n, nb=nb)
        tau = dt.to_numpy()
        np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
        got = da.to_numpy().reshape(n, m).T
        want = ref.reshape(n, m).T
        R, Rr = np.triu(got[:n]), np.triu(want[:n])
        np.testing.assert_allclose(R, Rr, rtol=0, atol=1e-10 * np.max(np.abs(Rr)))


def test_graded_spectrum_split_trailing_updates():
    """A = U diag(s) V^T with s geometric 1 -> 1e-8 (the tall-QR construction) with INT8 split
    trailing updates: reconstruction and orthogonality stay at FP64 level."""
    rng = np.random.default_rng(5)
    m, n = 4096, 1024
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.geomspace(1.0, 1e-8, n)
    A = (U * s) @ V.T
    res = {}
    for name, mode in (("fp64", native.MODE_FP64), ("split", native.MODE_SPLIT_INT8), ("crt", native.MODE_SPLIT_CRT)):
        da, d0, dt, _ = _factor(A, nb=128, mode=mode)
        res[name] = _errors(A, da, d0, dt, 128)
    print(res, flush=True)
    for name, (recon, orth) in res.items():
        assert recon < 1e-13 and orth < 1e-12, (name, recon, orth)


def test_wide_and_ragged_panels_against_oracle():
    """n > m (columns past min(m, n) take every panel's update, no reflector of their own) and a ragged
    last panel, on the caller's stream: R and tau as the oracle restatement, A = QR to FP64 level with Q
    formed by the oracle from the first min(m, n) columns."""
    rng = np.random.default_rng(91)
    stream = native.stream()
    for (m, n, nb) in [(300, 700, 64), (777, 520, 32), (130, 131, 128)]:
        A = rng.standard_normal((m, n))
        a_cm = np.asfortranarray(A).reshape(-1, order="F").copy()
        da = native.DeviceArray.from_numpy(a_cm)
        kk = min(m, n)
        dt = native.DeviceArray(kk)
        native.synchronize()
        native.dgeqrf_dev(m, n, da.ptr, m, dt.ptr, nb, native.MODE_FP64, stream)
        native.synchronize()
        F, tau = da.to_numpy().reshape(n, m).T, dt.to_numpy()
        ref = a_cm.copy()
        tau_ref = orc.qr_factor(ref, m, n, nb=nb)
        Fr = ref.reshape(n, m).T
        np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(np.triu(F[:kk]), np.triu(Fr[:kk]), rtol=0, atol=1e-10 * np.abs(Fr).max())
        f = np.asfortranarray(F[:, :kk]).reshape(-1, order="F").copy()
        Q = orc.qr_form_q(f, tau, m, kk).reshape(kk, m).T
        assert np.linalg.norm(A - Q @ np.triu(F[:kk])) / np.linalg.norm(A) < 1e-14, (m, n, nb)
        assert np.linalg.norm(Q.T @ Q - np.eye(kk)) < 1e-12, (m, n, nb)


def _oracle_factors(A, nb):
    m, n = A.shape
    ref = np.asfortranarray(A).reshape(-1, order="F").copy()
    tau_ref = orc.qr_factor(ref, m, n, nb=nb)
    return ref.reshape(n, m).T, tau_ref


def _cqr_panels(m, n, nb):
    """Panels the table's CholeskyQR2 bound sends down that path (at least qr_cqr_min_rows rows and twice
    the panel's width)."""
    lo = native.tune_get("qr_cqr_min_rows")
    kk = min(m, n)
    widths = [(k0, min(nb, kk - k0)) for k0 in range(0, kk, nb)]
    return sum(1 for k0, b in widths if lo > 0 and m - k0 >= lo and m - k0 >= 2 * b), len(widths)


def test_cholesky_qr2_panels_against_oracle():
    """Tall panels by CholeskyQR2 + Householder reconstruction: tau, R and the Householder vectors as the
    oracle restatement (computeHouseholder's conventions, HouseholderQR.java:219-291), A = QR and
    Q^T Q = I at FP64 level; with the path switched off (qr_cqr_min_rows 0) the column kernels give the
    same factors."""
    rng = np.random.default_rng(101)
    saved = native.tune_get("qr_cqr_min_rows")
    for (m, n, nb) in [(2048, 300, 128), (1024, 256, 64), (4096, 128, 128), (700, 100, 100), (1536, 1536, 128)]:
        A = rng.standard_normal((m, n))
        kk = min(m, n)
        Fr, tau_ref = _oracle_factors(A, nb)
        want, panels = _cqr_panels(m, n, nb)
        assert want > 0
        outs = []
        for lo in (saved, 0):
            native.tune_set("qr_cqr_min_rows", lo)
            try:
                da, d0, dt, st = _factor(A, nb=nb)
            finally:
                native.tune_set("qr_cqr_min_rows", saved)
            assert (st.tile_n, st.tile_k) == ((want, panels - want) if lo else (0, panels)), (m, n, nb, lo)
            F, tau = da.to_numpy().reshape(n, m).T, dt.to_numpy()
            np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(np.triu(F[:kk]), np.triu(Fr[:kk]), rtol=0, atol=1e-10 * np.abs(Fr).max())
            np.testing.assert_allclose(np.tril(F[:, :kk], -1), np.tril(Fr[:, :kk], -1), rtol=0, atol=1e-10)
            recon, orth = _errors(A, da, d0, dt, nb)
            assert recon < 1e-14 and orth < 1e-13, (m, n, nb, lo, recon, orth)
            outs.append(F)
        print(f"[qr] cqr {m}x{n} nb={nb}: {want} of {panels} panels, max |CQR - column| "
              f"{np.abs(outs[0] - outs[1]).max():.2e}", flush=True)


def test_narrow_cholesky_qr2_panels_against_oracle():
    """CholeskyQR2 panels at most 16 wide (every panel at nb = 16 or 8, and ragged 8- and 12-column last
    panels): their triangular inverses (X = R^-1, U^-1, T) are single 16 x 16 blocks, whose result one
    wavefront stores before the others read it. tau, R and the Householder vectors as the oracle's."""
    rng = np.random.default_rng(104)
    for (m, n, nb) in [(600, 160, 16), (512, 40, 8), (1000, 264, 128), (777, 520, 32), (900, 300, 96)]:
        A = rng.standard_normal((m, n))
        kk = min(m, n)
        Fr, tau_ref = _oracle_factors(A, nb)
        want, panels = _cqr_panels(m, n, nb)
        for rep in range(3):  # the same factors on every run
            da, d0, dt, st = _factor(A, nb=nb)
            assert st.tile_n == want and want > 0, (m, n, nb, st.tile_n, want)
            F, tau = da.to_numpy().reshape(n, m).T, dt.to_numpy()
            np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(np.triu(F[:kk]), np.triu(Fr[:kk]), rtol=0, atol=1e-10 * np.abs(Fr).max())
            np.testing.assert_allclose(np.tril(F[:, :kk], -1), np.tril(Fr[:, :kk], -1), rtol=0, atol=1e-10,
                                       err_msg=f"{(m, n, nb)} run {rep}")
            recon, orth = _errors(A, da, d0, dt, nb)
            assert recon < 1e-14 and orth < 1e-13, (m, n, nb, recon, orth)


def test_random_shapes_and_panel_widths_against_oracle():
    """Seeded shapes across every panel path (CholeskyQR2, one workgroup, column kernels), panel pairs and
    ragged last panels: m, n in [1, 1100] (tall, square and wide), nb in {8, ..., 128}; tau, R and the
    Householder vectors as the oracle's."""
    rng = np.random.default_rng(105)
    for _ in range(24):
        m = int(rng.integers(1, 1101))
        n = int(rng.integers(1, 601)) if rng.random() < 0.7 else int(rng.integers(1, 1101))
        nb = int(rng.choice([8, 16, 24, 32, 48, 64, 96, 128]))
        A = rng.standard_normal((m, n))
        kk = min(m, n)
        Fr, tau_ref = _oracle_factors(A, nb)
        da, d0, dt, st = _factor(A, nb=nb)
        F, tau = da.to_numpy().reshape(n, m).T, dt.to_numpy()
        tag = f"{(m, n, nb)} panels {st.tile_n}/{st.tile_k}"
        np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12, err_msg=tag)
        np.testing.assert_allclose(np.triu(F[:kk]), np.triu(Fr[:kk]), rtol=0, atol=1e-10 * np.abs(Fr).max(),
                                   err_msg=tag)
        np.testing.assert_allclose(np.tril(F[:, :kk], -1), np.tril(Fr[:, :kk], -1), rtol=0, atol=1e-10, err_msg=tag)
        if m >= n:
            recon, orth = _errors(A, da, d0, dt, nb)
            assert recon < 1e-14 and orth < 1e-13, (tag, recon, orth)


def test_register_panels_against_oracle_and_column_kernels():
    """Panels of at most qr_reg_max_m rows (and 128 columns) that CholeskyQR2 does not take are factored on
    one workgroup by the batched QR's register kernel: tau, R and the Householder vectors as the oracle's
    (and as the column kernels', qr_reg_max_m 0), A = QR and Q^T Q = I at FP64 level."""
    rng = np.random.default_rng(103)
    saved = native.tune_get("qr_reg_max_m")
    for (m, n, nb) in [(300, 300, 128), (192, 150, 64), (250, 200, 128), (160, 96, 96), (520, 300, 128),
                       (100, 300, 128), (50, 200, 64)]:  # the last two wider than tall
        A = rng.standard_normal((m, n))
        kk = min(m, n)
        Fr, tau_ref = _oracle_factors(A, nb)
        for reg in (saved, 0):
            native.tune_set("qr_reg_max_m", reg)
            try:
                da, d0, dt, st = _factor(A, nb=nb)
            finally:
                native.tune_set("qr_reg_max_m", saved)
            F, tau = da.to_numpy().reshape(n, m).T, dt.to_numpy()
            np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(np.triu(F[:kk]), np.triu(Fr[:kk]), rtol=0, atol=1e-10 * np.abs(Fr).max())
            np.testing.assert_allclose(np.tril(F[:, :kk], -1), np.tril(Fr[:, :kk], -1), rtol=0, atol=1e-10)
            if m >= n:  # (the error metrics take n <= m)
                recon, orth = _errors(A, da, d0, dt, nb)
                assert recon < 1e-14 and orth < 1e-13, (m, n, nb, reg, recon, orth)
            assert st.tile_n + st.tile_k == (kk + nb - 1) // nb, (m, n, nb, reg)


def test_cholesky_qr2_special_columns_and_fallback():
    """Upper-triangular panels (every column zero below its pivot: the reference's tau = 0 case for a
    non-negative pivot, tau = 2 with the row negated for a negative one) come out of the reconstruction
    exactly as the oracle's; a rank-deficient panel (a repeated column) and one of condition number 1e12
    fail CholeskyQR2's checks and run the column kernels, to the same FP64-level errors."""
    rng = np.random.default_rng(102)
    m, n, nb = 1024, 256, 128
    up = np.zeros((m, n))
    up[:n] = np.triu(rng.uniform(0.5, 1.0, (n, n)))
    signed = up.copy()
    signed[:n] *= rng.choice([-1.0, 1.0], n)[:, None]
    for name, A in (("upper", up), ("signed", signed)):
        da, d0, dt, st = _factor(A, nb=nb)
        assert (st.tile_n, st.tile_k) == (2, 0), name
        F, tau = da.to_numpy().reshape(n, m).T, dt.to_numpy()
        Fr, tau_ref = _oracle_factors(A, nb)
        assert np.array_equal(tau, tau_ref), name  # exactly 0 or 2
        # the reference returns R = A exactly; R2 R1 carries CholeskyQR's kappa * u rounding
        np.testing.assert_allclose(np.triu(F[:n]), np.triu(Fr[:n]), rtol=0, atol=1e-11, err_msg=name)
        assert not np.any(np.tril(F, -1)), name
    dup = rng.standard_normal((m, n))
    dup[:, 5] = dup[:, 3]
    ill = rng.standard_normal((m, n))
    u, _ = np.linalg.qr(rng.standard_normal((m, nb)))
    v, _ = np.linalg.qr(rng.standard_normal((nb, nb)))
    ill[:, :nb] = (u * np.geomspace(1.0, 1e-12, nb)) @ v.T  # condition number 1e12, not a column scaling
    for name, A in (("repeated column", dup), ("condition 1e12", ill)):
        da, d0, dt, st = _factor(A, nb=nb)
        assert st.tile_k >= 1 and st.tile_n + st.tile_k == 2, (name, st.tile_n, st.tile_k)
        recon, orth = _errors(A, da, d0, dt, nb)
        assert recon < 1e-14 and orth < 1e-13, (name, recon, orth)


def _batch(rng, count, lo=16, hi=192):
    ms = rng.integers(lo, hi + 1, count)
    ns = np.maximum(1, ms // 2)
    mats = [rng.standard_normal((int(m), int(n))) for m, n in zip(ms, ns)]
    flat = np.concatenate([np.asfortranarray(M).reshape(-1, order="F") for M in mats])
    offs = np.concatenate([[0], np.cumsum([M.size for M in mats])[:-1]])
    tau_offs = np.concatenate([[0], np.cumsum([min(M.shape) for M in mats])[:-1]])
    return mats, flat, offs, tau_offs


def test_batched_small_qr_against_oracle_and_graph_replay():
    rng = np.random.default_rng(77)
    mats, flat, offs, toffs = _batch(rng, 600)
    da = native.DeviceArray.from_numpy(flat)
    dt = native.DeviceArray(int(sum(min(M.shape) for M in mats)))
    descs = [native.QrDesc(M.shape[0], M.shape[1], M.shape[0], da.at(int(o)), dt.at(int(t)))
             for M, o, t in zip(mats, offs, toffs)]
    st = native.dgeqrf_batched_dev(descs)
    assert st.kernel_name == "qr"
    got = da.to_numpy()
    taus = dt.to_numpy()
    for i, M in enumerate(mats):
        m, n = M.shape
        f = got[offs[i]:offs[i] + m * n].copy()
        tau = taus[toffs[i]:toffs[i] + n]
        ref = np.asfortranarray(M).reshape(-1, order="F").copy()
        tau_ref = orc.qr_factor(ref, m, n)
        np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
        Q = orc.qr_form_q(f, tau, m, n).reshape(n, m).T
        R = np.triu(f.reshape(n, m).T[:n])
        assert np.linalg.norm(M - Q @ R) / np.linalg.norm(M) < 1e-13
        assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13
    # graph replays give the eager result again, bit for bit
    plan = native.QrBatchedPlan(descs)
    for _ in range(3):
        native.check(native.load().jlc_memcpy_h2d(da.ptr, flat.ctypes.data, flat.size * 8))
        native.synchronize()
        plan.replay()
        native.synchronize()
        assert np.array_equal(da.to_numpy(), got)
    plan.destroy()


def test_batched_rejects_oversized_matrix():
    d = native.DeviceArray(8)
    for m, n in ((300, 100), (500, 41), (2, 257)):  # m*n past the LDS tile (20220 entries), n past 256
        with pytest.raises(native.JlcError) as ei:
            native.QrBatchedPlan([native.QrDesc(m, n, m, d.ptr, d.ptr)])
        assert ei.value.status == native.JLC_ERR_INVALID_ARG


def test_batched_register_kernel_extreme_scales_against_oracle():
    # Matrices of the register kernel's sizes scaled near the top of the FP64 range (column maxima above
    # 2^1000, where the reflector scales by 2^-100 before taking 1 / max) and near the bottom (1 / max still
    # normal); the oracle forms x / max by division.
    rng = np.random.default_rng(79)
    mats = []
    for (m, n), scale in (((40, 20), 1e305), ((150, 75), 3e305), ((192, 96), 1e305), ((64, 32), 1e-300),
                          ((100, 50), 2e-290), ((33, 16), 1.0)):
        mats.append(rng.uniform(-1.0, 1.0, (m, n)) * scale)
    mats[5][:, 3] *= 1e303  # one column far above the others
    flat = np.concatenate([np.asfortranarray(M).reshape(-1, order="F") for M in mats])
    offs = np.concatenate([[0], np.cumsum([M.size for M in mats])[:-1]])
    toffs = np.concatenate([[0], np.cumsum([min(M.shape) for M in mats])[:-1]])
    da = native.DeviceArray.from_numpy(flat)
    dt = native.DeviceArray(int(sum(min(M.shape) for M in mats)))
    descs = [native.QrDesc(M.shape[0], M.shape[1], M.shape[0], da.at(int(o)), dt.at(int(t)))
             for M, o, t in zip(mats, offs, toffs)]
    native.dgeqrf_batched_dev(descs)
    got, taus = da.to_numpy(), dt.to_numpy()
    for i, M in enumerate(mats):
        m, n = M.shape
        f = got[offs[i]:offs[i] + m * n].copy()
        tau = taus[toffs[i]:toffs[i] + n]
        assert np.all(np.isfinite(f)) and np.all(np.isfinite(tau))
        ref = np.asfortranarray(M).reshape(-1, order="F").copy()
        tau_ref = orc.qr_factor(ref, m, n)
        np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)
        Rg, Rr = np.triu(f.reshape(n, m).T[:n]), np.triu(ref.reshape(n, m).T[:n])
        np.testing.assert_allclose(Rg, Rr, rtol=1e-9, atol=1e-12 * np.abs(Rr).max())
        Q = orc.qr_form_q(f, tau, m, n).reshape(n, m).T
        assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13
        # A = QR compared after scaling both by the matrix's norm (products of 1e305 entries would overflow)
        s = 1.0 / np.abs(M).max()
        assert np.linalg.norm(M * s - Q @ (Rg * s)) / np.linalg.norm(M * s) < 1e-13


def test_batched_largest_tiles_and_special_columns():
    # The largest accepted tiles (m*n = 20220 and n = 256), beside matrices whose first column is zero,
    # zero below a negative pivot, or below the FP64 minimum.
    rng = np.random.default_rng(78)
    mats = [rng.standard_normal((674, 30)), rng.standard_normal((78, 256)), rng.standard_normal((30, 12)),
            rng.standard_normal((25, 10)), rng.standard_normal((20, 8))]
    # (the first column meets no earlier reflector, so each case happens as built)
    mats[2][:, 0] = 0.0                 # tau = 0
    mats[3][:, 0] = 0.0
    mats[3][0, 0] = -1.5                # zero below a negative pivot: tau = 2, R(0,0) = 1.5
    mats[4][:, 0] = 1e-310              # max below the FP64 minimum: tau = 0
    flat = np.concatenate([np.asfortranarray(M).reshape(-1, order="F") for M in mats])
    offs = np.concatenate([[0], np.cumsum([M.size for M in mats])[:-1]])
    toffs = np.concatenate([[0], np.cumsum([min(M.shape) for M in mats])[:-1]])
    da = native.DeviceArray.from_numpy(flat)
    dt = native.DeviceArray(int(sum(min(M.shape) for M in mats)))
    descs = [native.QrDesc(M.shape[0], M.shape[1], M.shape[0], da.at(int(o)), dt.at(int(t)))
             for M, o, t in zip(mats, offs, toffs)]
    native.dgeqrf_batched_dev(descs)
    got, taus = da.to_numpy(), dt.to_numpy()
    for i, M in enumerate(mats):
        m, n = M.shape
        kk = min(m, n)
        f = got[offs[i]:offs[i] + m * n].copy()
        tau = taus[toffs[i]:toffs[i] + kk]
        ref = np.asfortranarray(M).reshape(-1, order="F").copy()
        tau_ref = orc.qr_factor(ref, m, n)
        np.testing.assert_allclose(tau, tau_ref[:kk], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(np.triu(f.reshape(n, m).T[:kk]), np.triu(ref.reshape(n, m).T[:kk]),
                                   rtol=1e-9, atol=1e-12 * np.abs(M).max())
