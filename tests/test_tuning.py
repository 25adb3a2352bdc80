"""Tuning-table ABI and its use by dispatch (host logic: no GPU needed)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

from jlc_amd import native
from jlc_amd.compute import DispatchPolicy, GemmDispatch

ROOT = Path(__file__).resolve().parent.parent


def test_table_round_trip(tmp_path):
    old = {k: native.tune_get(k) for k in ("hip_min_flops", "splitk_tiles_per_cu", "splitk_min_kslice",
                                           "grouped_waves_per_simd", "grouped_xcds")}
    try:
        native.tune_set("hip_min_flops", 12345678.0)
        native.tune_set("splitk_tiles_per_cu", 3)
        p = tmp_path / "t.tune"
        native.tune_write(str(p))
        native.tune_set("hip_min_flops", 1.0)
        native.tune_load(str(p))
        assert native.tune_get("hip_min_flops") == 12345678.0
        assert native.tune_get("splitk_tiles_per_cu") == 3
        bad = tmp_path / "bad.tune"
        bad.write_text("no_such_key 3\n")
        with pytest.raises(native.JlcError):
            native.tune_load(str(bad))
        with pytest.raises(native.JlcError):
            native.tune_get("no_such_key")
    finally:
        for k, v in old.items():
            native.tune_set(k, v)


def test_dispatch_reads_the_table():
    old = native.tune_get("hip_min_flops")
    try:
        native.tune_set("hip_min_flops", 1000.0)
        pol = DispatchPolicy.defaultPolicy()
        assert pol.hip_min_flops == 1000
        assert GemmDispatch.selectKernel(8, 8, 8, True, 1, pol) is GemmDispatch.Kernel.HIP
        native.tune_set("hip_min_flops", 1e12)
        pol = DispatchPolicy.defaultPolicy()
        assert GemmDispatch.selectKernel(512, 512, 512, True, 8, pol) is GemmDispatch.Kernel.PARALLEL_MICRO
    finally:
        native.tune_set("hip_min_flops", old)


def test_committed_table_loads_in_a_fresh_process():
    table = native.TUNE_TABLE
    if not table.exists():
        pytest.skip("no tuning table generated yet")
    code = f"import sys; sys.path.insert(0, {str(ROOT)!r}); from jlc_amd import native; print(native.tune_get('hip_min_flops'))"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert float(out.stdout.strip()) > 0
