"""bench.py's contract pieces that run without a GPU: the one-GPU guard and the
HBM traffic figures it reads from the committed rocprofv3 summaries."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_traffic_summaries_parse():
    import bench
    # kernel tag -> (lower, upper) plausible HBM bytes per launch
    cases = {"crt_gemm_i8_16384": (1e10, 1e12), "dgemm_f64_mfma_16384": (1e10, 1e12),
             "dgemv_32768": (8e9, 9e9), "grouped_f64_100k": (7.9e9, 2e10),
             "dgemm_f64_mfma_qr_update_32768x4096": (1.0e9, 2e9), "qr_batched_20000": (2.1e9, 3e9),
             "dgemm_f64_mfma_8192": (1.6e9, 1e11)}
    for tag, (lo, hi) in cases.items():
        v = bench.read_traffic(tag)
        assert v is not None and lo <= v <= hi, (tag, v)
    assert bench.read_traffic("no_such_kernel") is None


def _form(name, tflops, kernel, slices, rel):
    return {"form": name, "tflops": tflops, "kernel": kernel, "slices": slices, "error": {"max_rel_err": rel}}


def test_value_form_skips_fp64_fallbacks():
    """A split form answered by the FP64 kernel (automatic mode, spread too wide) has FP64's error, so
    it would pass the error rule; it must never be picked, however the timing noise falls."""
    import bench
    fp64 = _form("fp64_mfma", 67.0, "mfma_f64", 0, 1e-12)
    fallback = _form("split_bf16", 67.4, "mfma_f64", 0, 1e-12)
    assert not bench.split_ran(fallback)
    assert bench.choose_value_form(fp64, [fallback]) is fp64
    crt = _form("split_crt", 120.0, "split_crt", 15, 5e-13)
    assert bench.split_ran(crt)
    assert bench.choose_value_form(fp64, [fallback, crt]) is crt
    # a faster split form with a larger error than FP64's is not the value form
    loose = _form("split_int8", 150.0, "split_int8", 5, 2e-12)
    assert bench.choose_value_form(fp64, [loose, crt]) is crt
    assert bench.choose_value_form(fp64, [loose], check_error=False) is loose
    # slower split forms lose to FP64
    slow = _form("split_int8", 50.0, "split_int8", 9, 1e-13)
    assert bench.choose_value_form(fp64, [slow]) is fp64
