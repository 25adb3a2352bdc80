"""The automatic slice / moduli choosers over the benchmark's exponent spreads, at a reduced size
(512 x 512 x 2048, a = (u - 0.5) exp(phi g) for phi in {0.5, 2, 8}).

For each phi and split form the test pins which kernel ran and with how many slices / moduli, and that
a form which ran its own kernels has a max relative error against double-double dot products, on a
128 x 128 sample, at most the FP64 MFMA kernel's on the same inputs, so that automatic mode never hands
back a less accurate result than FP64. A form that cannot hold the spread runs the FP64 kernel and
returns its result bit for bit. The FP64 reference is the unsplit kernel (tests/fp64_ref.py). The
CRT pick is also checked against its defining rule: the fewest moduli that hold the chosen norm bits,
with the leftover capacity spent on bits.
"""
import numpy as np
import pytest

from jlc_amd import native
from oracle import oracle as orc
from tests.fp64_ref import unsplit_fp64

pytestmark = pytest.mark.gpu

M, N, K = 512, 512, 2048

# phi -> form -> (kernel, slices or moduli); ("mfma_f64", 0) is the FP64 kernel serving the call.
PICKS = {
    0.5: {"split_crt": ("split_crt", 15), "split_int8": ("split_int8", 9), "split_bf16": ("split_bf16", 8)},
    2.0: {"split_crt": ("split_crt", 16), "split_int8": ("mfma_f64", 0), "split_bf16": ("mfma_f64", 0)},
    8.0: {"split_crt": ("mfma_f64", 0), "split_int8": ("mfma_f64", 0), "split_bf16": ("mfma_f64", 0)},
}
MODES = {"split_crt": native.MODE_SPLIT_CRT, "split_int8": native.MODE_SPLIT_INT8,
         "split_bf16": native.MODE_SPLIT_BF16}


def operand(rng, rows, cols, phi):
    return (rng.random(rows * cols) - 0.5) * np.exp(phi * rng.standard_normal(rows * cols))


@pytest.mark.parametrize("phi", sorted(PICKS))
def test_chooser_picks_and_accuracy(phi):
    rng = np.random.default_rng(int(phi * 1000) + 11)
    a, b = operand(rng, M, K, phi), operand(rng, K, N, phi)
    da, db, dc = native.DeviceArray.from_numpy(a), native.DeviceArray.from_numpy(b), native.DeviceArray(M * N)
    ii, jj = np.meshgrid(np.arange(0, M, 4), np.arange(0, N, 4), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    hi, lo, _ = orc.dd_entries(native.ROW, native.ROW, K, a, K, b, N, ii, jj)
    exact = hi + lo
    nz = exact != 0.0

    def run(mode):
        with unsplit_fp64():
            st = native.dgemm_dev(native.ROW, native.ROW, M, N, K, 1.0, da.ptr, K, db.ptr, N, 0.0, dc.ptr, N, mode)
        c = dc.to_numpy()
        got = c.reshape(M, N)[ii, jj]
        return st, c, float(np.max(np.abs(got[nz] - exact[nz]) / np.abs(exact[nz])))

    _, c64, rel64 = run(native.MODE_FP64)
    picks = {}
    for form, mode in MODES.items():
        st, c, rel = run(mode)
        picks[form] = (st.kernel_name, st.slices)
        print(f"phi={phi} {form}: {st.kernel_name} ({st.slices}) rel {rel:.3e} fp64 {rel64:.3e}", flush=True)
        if st.kernel_name == "mfma_f64":
            assert np.array_equal(c, c64)
        else:
            assert rel <= rel64, (phi, form, rel, rel64)
        if st.kernel_name == "split_crt":
            assert orc.crt_fits(st.bits_a, st.bits_b, K, st.slices)
            assert not orc.crt_fits(st.bits_a, st.bits_b, K, st.slices - 1)
            more = (st.bits_a + 1, st.bits_b) if st.bits_a <= st.bits_b else (st.bits_a, st.bits_b + 1)
            assert max(more) > orc.CRT_MAX_BITS or not orc.crt_fits(*more, K, st.slices)
    assert picks == PICKS[phi], picks
