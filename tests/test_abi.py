"""The C-ABI library loads, exports every declared entry point, and rejects
bad arguments before touching a device (no compute calls: runs without a GPU)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from jlc_amd import native


def test_library_loads_and_exports_every_header_symbol():
    lib = native.load()
    names = native.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    unbound = [n for n in names if n not in native._SIGNATURES]
    assert not unbound, unbound


def test_library_has_no_torch_or_python_dependency():
    out = subprocess.run(["ldd", str(native.LIB_PATH)], capture_output=True, text=True).stdout
    libs = [ln.split()[0] for ln in out.splitlines() if ln.strip()]
    assert not [x for x in libs if "torch" in x or "python" in x or x.startswith("libc10")], libs


def test_version_string():
    assert b"gfx950" in native.load().jlc_version()


def _rc(*args):
    st = native.Stats()
    return native.load().jlc_dgemm(*args, C.byref(st))


def test_invalid_arguments_rejected_without_device():
    a = np.zeros(16)
    p = a.ctypes.data
    cases = [
        (2, 0, 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2, 0),     # bad layout_a
        (0, 5, 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2, 0),     # bad layout_b
        (0, 0, -1, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2, 0),    # negative m
        (0, 0, 2, 2, 2, 1.0, p, 1, p, 2, 0.0, p, 2, 0),     # lda < k
        (0, 0, 2, 2, 2, 1.0, p, 2, p, 1, 0.0, p, 2, 0),     # ldb < n
        (0, 0, 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 1, 0),     # ldc < n
        (0, 0, 2, 2, 2, 1.0, None, 2, p, 2, 0.0, p, 2, 0),  # null A
        (0, 0, 2, 2, 2, 1.0, p, 2, p, 2, 0.0, None, 2, 0),  # null C
        (1, 0, 3, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2, 0),     # COL A: lda < m
        (0, 0, 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2, 77),    # unknown mode
    ]
    for cs in cases:
        assert _rc(*cs) == native.JLC_ERR_INVALID_ARG, cs
        assert native.last_error(), cs


def test_empty_outputs_are_ok_without_device():
    # m == 0 or n == 0: nothing to do (GemmParityTest.java:118-124)
    assert _rc(0, 0, 0, 2, 3, 1.0, None, 3, None, 2, 0.0, None, 2, 0) == native.JLC_OK
    assert _rc(0, 0, 2, 0, 0, 1.0, None, 1, None, 1, 0.0, None, 1, 0) == native.JLC_OK
    st = native.Stats()
    assert native.load().jlc_dgemv(0, 5, 1.0, None, 5, None, 1, 0.0, None, 1, C.byref(st)) == native.JLC_OK


def test_disabled_by_environment():
    code = ("import sys; sys.path.insert(0, %r); from jlc_amd import native; "
            "rc, n, _ = native.probe(); print(rc)") % str(native.LIB_PATH.parent.parent.parent)
    env = dict(os.environ, JLC_HIP_ENABLED="0")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() == str(native.JLC_ERR_DISABLED), out.stderr


def test_mode_encoding():
    assert native.mode_slices(native.MODE_SPLIT_INT8, 7) == 0x701
