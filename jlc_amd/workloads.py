"""Synthetic workloads of the measured configurations (device-resident).

* batched_shapes: the grouped-GEMM configuration — M, N, K each drawn
  log-uniformly over [2, 256], a fraction of the problems forced to
  2 x 2 x 2, 3 x 3 x 3 or 4 x 4 x 4 cubes, alpha and beta uniform on [-2, 2].
* BatchedWorkload: all operands of such a batch packed into three device
  buffers (entries uniform on [-1, 1]) and the matching descriptors.
"""
from __future__ import annotations

import numpy as np

from . import native


def batched_shapes(count: int, seed: int, lo: int = 2, hi: int = 256, cube_frac: float = 0.1):
    rng = np.random.default_rng(seed)
    dims = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))
    dims = np.clip(np.rint(dims), lo, hi).astype(np.int64)
    cubes = rng.random(count) < cube_frac
    side = rng.integers(2, 5, size=count)
    dims[cubes] = side[cubes, None]
    alpha = rng.uniform(-2.0, 2.0, count)
    beta = rng.uniform(-2.0, 2.0, count)
    return dims, alpha, beta


class BatchedWorkload:
    """Row-major operands of every problem packed back to back in device memory."""

    def __init__(self, dims: np.ndarray, alpha: np.ndarray, beta: np.ndarray, seed: int = 0,
                 layouts: np.ndarray | None = None):
        self.dims = dims
        self.alpha = alpha
        self.beta = beta
        self.count = len(dims)
        m, n, k = dims[:, 0], dims[:, 1], dims[:, 2]
        self.layouts = layouts if layouts is not None else np.zeros((self.count, 2), dtype=np.int64)
        a_sz, b_sz, c_sz = m * k, k * n, m * n
        # Start every operand on a 16-byte boundary.
        pad = lambda x: (x + 1) // 2 * 2
        self.a_off = np.concatenate([[0], np.cumsum(pad(a_sz))[:-1]])
        self.b_off = np.concatenate([[0], np.cumsum(pad(b_sz))[:-1]])
        self.c_off = np.concatenate([[0], np.cumsum(pad(c_sz))[:-1]])
        rng = np.random.default_rng(seed)
        self.a_host = rng.uniform(-1.0, 1.0, int(self.a_off[-1] + pad(a_sz[-1])))
        self.b_host = rng.uniform(-1.0, 1.0, int(self.b_off[-1] + pad(b_sz[-1])))
        self.c_host = rng.uniform(-1.0, 1.0, int(self.c_off[-1] + pad(c_sz[-1])))
        self.da = native.DeviceArray.from_numpy(self.a_host)
        self.db = native.DeviceArray.from_numpy(self.b_host)
        self.dc = native.DeviceArray.from_numpy(self.c_host)
        self.bytes = int(8 * np.sum(a_sz + b_sz + c_sz * (1 + (beta != 0.0))))
        self.flops = int(2 * np.sum(m * n * k))

    def lds(self, i: int) -> tuple[int, int]:
        m, n, k = (int(x) for x in self.dims[i])
        la, lb = (int(x) for x in self.layouts[i])
        return (k if la == native.ROW else m), (n if lb == native.ROW else k)

    def descs(self) -> list:
        out = []
        for i in range(self.count):
            m, n, k = (int(x) for x in self.dims[i])
            lda, ldb = self.lds(i)
            la, lb = (int(x) for x in self.layouts[i])
            out.append(native.GemmDesc(la, lb, m, n, k, 0, float(self.alpha[i]), float(self.beta[i]),
                                       self.da.at(int(self.a_off[i])), self.db.at(int(self.b_off[i])),
                                       self.dc.at(int(self.c_off[i])), lda, ldb, n))
        return out

    def reset_c(self) -> None:
        native.check(native.load().jlc_memcpy_h2d(self.dc.ptr, self.c_host.ctypes.data, self.c_host.size * 8))

    def problem(self, i: int, c_now: np.ndarray | None = None):
        """(A, B, C0, C_now) flat views of problem i in the host copies (C_now from c_now if given)."""
        m, n, k = (int(x) for x in self.dims[i])
        A = self.a_host[self.a_off[i]:self.a_off[i] + m * k]
        B = self.b_host[self.b_off[i]:self.b_off[i] + k * n]
        C0 = self.c_host[self.c_off[i]:self.c_off[i] + m * n]
        Cn = None if c_now is None else c_now[self.c_off[i]:self.c_off[i] + m * n]
        return A, B, C0, Cn
