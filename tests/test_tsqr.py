"""Tall-skinny QR: TallSkinnyQR.decompose / CommunicationAvoidingQR.decompose
(src/main/java/net/faulj/decomposition/qr/TallSkinnyQR.java:22-90, CommunicationAvoidingQR.java:18-20).

The oracle's restatement (oracle_tsqr: row blocks of max(n, 64) rows, thin QRs by the HouseholderQR
restatement, a QR of the stacked R factors, Q blocks as block-Q times stack-Q) is pinned on the CPU by
what the mathematics fixes: A = QR to rounding, Q^T Q = I, R upper trapezoidal, and |R| equal row by
row to |R| of LAPACK's QR of the same A (a full-rank QR is unique up to the signs of R's rows). The
reference's own acceptance test is restated with its sizes and tolerances
(DecompositionTests.java:583-615, :761-793: orthonormal columns within 10 tol, reconstruction within
50 tol, tol = getTolerance(max(m, n))). The GPU tests compare jlc_dtsqr[_dev] with the restatement
(the same algorithm and reflector convention, so Q and R agree to rounding, signs included) from the
reference's sizes up to 65536 x 32 and 32768 x 128.
"""
import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc

REFERENCE_TSQR = [(60, 10), (80, 5), (40, 12)]        # DecompositionTests.java:588-592
REFERENCE_CAQR = [(72, 9), (55, 7), (48, 16)]         # DecompositionTests.java:606-610


def _tolerance(n: int) -> float:
    """DecompositionTests.getTolerance (DecompositionTests.java:183-188)."""
    if n <= 10:
        return 1e-10 * np.sqrt(n)
    if n <= 30:
        return 1e-8 * np.sqrt(n)
    if n <= 100:
        return 1e-6 * np.sqrt(n)
    return 1e-5 * np.sqrt(n)


def _oracle(A):
    m, n = A.shape
    k = min(m, n)
    q, r = orc.tsqr(np.asfortranarray(A).reshape(-1, order="F").copy(), m, n)
    return q.reshape(k, m).T, r.reshape(n, k).T


@pytest.mark.parametrize("m,n", REFERENCE_TSQR + REFERENCE_CAQR + [(300, 70), (129, 64), (1000, 3), (5, 9),
                                                                   (200, 200), (65, 64)])
def test_oracle_tsqr_against_the_mathematics(m, n):
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    Q, R = _oracle(A)
    k = min(m, n)
    assert Q.shape == (m, k) and R.shape == (k, n)
    assert np.array_equal(R, np.triu(R))
    assert np.linalg.norm(A - Q @ R) <= 1e-14 * np.sqrt(m * n) * np.linalg.norm(A)
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) <= 1e-14 * np.sqrt(m) * k
    R_lapack = np.linalg.qr(A, mode="r")[:k]
    assert np.max(np.abs(np.abs(R) - np.abs(R_lapack))) <= 1e-13 * np.linalg.norm(A)
    tol = _tolerance(max(m, n))
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 10 * tol and np.linalg.norm(A - Q @ R) < 50 * tol


# ------------------------------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("m,n", REFERENCE_TSQR + REFERENCE_CAQR + [(300, 70), (1000, 3), (5, 9), (200, 200)])
def test_tsqr_through_the_mirror_like_the_reference_test(m, n):
    """testTallSkinnyQR / testCommunicationAvoidingQR (DecompositionTests.java:761-793) on the GPU path,
    and the factors against the restatement."""
    from jlc_amd.compute import CommunicationAvoidingQR, Matrix, TallSkinnyQR
    rng = np.random.default_rng(2900 + m * 10 + n)
    A = Matrix.wrap(rng.uniform(-1, 1, m * n), m, n)
    k = min(m, n)
    for decompose in (TallSkinnyQR.decompose, CommunicationAvoidingQR.decompose):
        res = decompose(A)
        Q, R = res.getQ(), res.getR()
        assert Q.getColumnCount() == k and R.getRowCount() == k
        tol = _tolerance(max(m, n))
        Qn, Rn = Q.to_numpy(), R.to_numpy()
        assert np.linalg.norm(Qn.T @ Qn - np.eye(k)) < 10 * tol
        QR = Q.multiply(R) if n > 64 and 2 * m * n * k >= 14155776 else Matrix.wrap((Qn @ Rn).reshape(-1), m, n)
        assert np.linalg.norm(QR.to_numpy() - A.to_numpy()) < 50 * tol
        Qo, Ro = _oracle(A.to_numpy())
        assert np.max(np.abs(Rn - Ro)) <= 1e-12 * np.max(np.abs(Ro))
        assert np.max(np.abs(Qn - Qo)) <= 1e-12


@pytest.mark.gpu
def test_tsqr_degenerate_shapes_follow_the_reference():
    from jlc_amd.compute import IllegalArgumentException, Matrix, TallSkinnyQR
    A = Matrix(4, 0)
    res = TallSkinnyQR.decompose(A)  # m > 0, n == 0: Q = I_m, R = A's copy (TallSkinnyQR.java:31-33)
    assert np.array_equal(res.getQ().to_numpy(), np.eye(4)) and res.getR().getColumnCount() == 0
    with pytest.raises(IllegalArgumentException):
        TallSkinnyQR.decompose(Matrix(0, 3))  # Matrix.Identity(0) rejects a zero size
    with pytest.raises(IllegalArgumentException):
        TallSkinnyQR.decompose(None)


@pytest.mark.gpu
@pytest.mark.parametrize("m,n", [(65536, 32), (32768, 128), (8192, 300), (70000, 17)])
def test_tsqr_device_large_against_restatement(m, n):
    """Batched block QRs and batched Q application (n <= 139), and the per-block blocked path (n = 300)."""
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    a_cm = np.asfortranarray(A).reshape(-1, order="F").copy()
    da = native.DeviceArray.from_numpy(a_cm)
    k = min(m, n)
    dq, dr = native.DeviceArray(m * k), native.DeviceArray(k * n)
    st = native.dtsqr_dev(m, n, da.ptr, m, dq.ptr, m, dr.ptr, k)
    print(f"tsqr {m}x{n}: {st.kernel_ms:.2f} ms", flush=True)
    assert np.array_equal(da.to_numpy(), a_cm)  # A is not modified
    Q = dq.to_numpy().reshape(k, m).T
    R = dr.to_numpy().reshape(n, k).T
    assert np.array_equal(R, np.triu(R))
    assert np.linalg.norm(A - Q @ R) / np.linalg.norm(A) < 1e-14
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-13
    Qo, Ro = _oracle(A)
    assert np.max(np.abs(R - Ro)) <= 1e-12 * np.max(np.abs(Ro))
    assert np.max(np.abs(Q - Qo)) <= 1e-11
