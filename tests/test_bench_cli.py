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
             "panel_col_qr_32768x4096": (3.1e7, 1e8), "qr_batched_20000": (2.1e9, 3e9),
             "dgemm_f64_mfma_8192": (1.6e9, 1e11)}
    for tag, (lo, hi) in cases.items():
        v = bench.read_traffic(tag)
        assert v is not None and lo <= v <= hi, (tag, v)
    assert bench.read_traffic("no_such_kernel") is None
