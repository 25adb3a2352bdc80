"""Writes gemm_kat.json: the known-answer cases the reference's own tests hold.

Each case is data transcribed from a reference test (cited), expressed as
inputs + expected outputs under this engine's BLAS semantics (beta == 0 never
reads C; alpha == 0 never reads A or B). Non-finite values are encoded as the
strings "nan" / "inf" / "-inf".

Run: python tests/golden/make_kat.py
"""
import json
from pathlib import Path

NAN, INF = "nan", "inf"

CASES = [
    {
        "name": "beta0_clears_nan_c",
        "source": "src/test/java/net/faulj/compute/GemmNumericEdgeTest.java:32-48 (BLAS3Kernels path); "
                  "also BatchProcessingTest.java:19-32",
        "m": 2, "k": 2, "n": 2, "alpha": 1.0, "beta": 0.0,
        "a": [1, 2, 3, 4], "b": [5, 6, 7, 8], "c": [NAN, NAN, NAN, NAN],
        "expected": [19, 22, 43, 50],
    },
    {
        "name": "alpha0_scales_c",
        "source": "src/test/java/net/faulj/compute/GemmNumericEdgeTest.java:50-64",
        "m": 2, "k": 2, "n": 2, "alpha": 0.0, "beta": 2.5,
        "a": [3, -1, 5, 2], "b": [8, -4, 7, 9], "c": [2, -6, 10, -14],
        "expected": [5, -15, 25, -35],
    },
    {
        "name": "alpha0_ignores_nan_inputs",
        "source": "src/test/java/net/faulj/compute/GemmNumericEdgeTest.java:87-96",
        "m": 1, "k": 1, "n": 1, "alpha": 0.0, "beta": -0.5,
        "a": [NAN], "b": [NAN], "c": [4.0],
        "expected": [-2.0],
    },
    {
        "name": "zero_times_inf_is_nan",
        "source": "src/test/java/net/faulj/compute/GemmNumericEdgeTest.java:77-85 (OptimizedBLAS3 path; "
                  "the BLAS3Kernels naive zero-skip of :66-75 is a CPU-only regime)",
        "m": 1, "k": 1, "n": 1, "alpha": 1.0, "beta": 1.0,
        "a": [0.0], "b": [INF], "c": [7.0],
        "expected": [NAN],
    },
]


def main() -> None:
    out = Path(__file__).with_name("gemm_kat.json")
    out.write_text(json.dumps({"cases": CASES}, indent=1) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
