"""Runs the C++ host mirror's parity program (tests/cpp/gemm_parity_test.cpp).

The program mirrors GemmParityTest / GemmStridedParityTest / GemmOffHeapParityTest /
GemmNumericEdgeTest / DispatchPolicyTest (src/test/java/net/faulj/compute/) through
the C++ classes in jlc_amd/host/ over libjlc_hip's C ABI. --cpu runs the cases that
need no device (dispatch, validation, degenerate shapes, off-heap sync).
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "build" / "gemm_parity_test"


def _binary() -> Path:
    if not BIN.exists():
        from jlc_amd import build as b
        if not (b.LIBDIR / b.LIBNAME).exists():
            b.build_hip(verbose=False)
        if not (ROOT / "oracle" / "build" / "liboracle.so").exists():
            b.build_oracle(verbose=False)
        b.build_host(verbose=False)
        b.build_cpp_tests(verbose=False)
    return BIN


def _run(args, timeout):
    r = subprocess.run([str(_binary()), *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
    return r.stdout


def test_cpp_mirror_cpu_cases():
    out = _run(["--cpu"], 60)
    assert "6 passed" in out


@pytest.mark.gpu
def test_cpp_mirror_parity_on_gpu():
    out = _run([], 110)
    assert "[ FAIL ]" not in out
    assert "19 passed" in out
