"""Autotuner for libjlc_hip on MI355X: measures candidate settings on the GPU
and writes the table (jlc_amd/tuning/mi355x.tune) that dispatch reads.

Tuned entries:
* splitk_tiles_per_cu / splitk_min_kslice: when a GEMM has few output tiles for
  its K (the Gram matrix and W = C^T V of blocked QR, tall-skinny products, and
  small products that leave most CUs idle), K is cut into slices; candidates are
  timed on such shapes and on square ones (which must not regress).
* grouped_waves_per_simd: width of the persistent grouped launch; grouped_small_tile, the m (n)
  up to which a problem runs 16-row (16-column) MFMA tiles instead of 32; grouped_chunk_tiles,
  how many tiles of one problem go to one XCD together; all timed on a batch of uneven problems.
* qr_wave_max_lds / qr_block_threads: the batched QR's size-class boundary (in LDS doubles per
  matrix) up to which a matrix is factored by one wavefront (cheap barriers, more matrices per CU),
  and the threads per matrix above it (256, 512 or 1024: more lanes on each column step); and
  qr_reg_max_m, the largest m (<= 192, with n <= 128) held in registers instead of LDS; timed on
  a batch with the config's shape distribution (m uniform on [16, 192], n = m / 2).
* qr_cqr_min_rows: the panel height from which blocked QR factors a 128-wide panel by CholeskyQR2 with
  Householder reconstruction (a fixed handful of launches) instead of one launch per column; single
  panels of each candidate height are timed both ways and the bound is the smallest height from which
  CholeskyQR2 stays faster.
* hip_min_flops: the offload gate of the host-side dispatch (the HIP analogue
  of GemmDispatch.CUDA_THRESHOLD, src/main/java/net/faulj/compute/GemmDispatch.java:32).
  The end-to-end host-pointer GEMM (staging included) is timed per size and
  compared with the CPU cost model of GemmDispatch.preferCuda
  (GemmDispatch.java:155-178: 100 GFLOP/s unless --cpu-gflops says otherwise);
  the gate is the smallest flop count from which the GPU stays faster.

Usage: python -m jlc_amd.tune [--out PATH] [--cpu-gflops G] [--quick]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import native
from .workloads import BatchedWorkload, batched_shapes


def _time_dev(fn, iters: int) -> float:
    fn()
    native.synchronize()
    t = native.Timer(native.stream())
    t.start()
    for _ in range(iters):
        fn()
    return t.stop() / iters


def tune_splitk(quick: bool) -> dict:
    # tall-skinny and long-K products (the Gram matrix and W = C^T V of blocked QR), square ones that must
    # not regress, and small ones where a few 128 x 128 tiles leave most CUs idle
    shapes = [(128, 128, 32768), (3968, 128, 32768), (1024, 128, 16384), (2048, 2048, 16384), (8192, 8192, 4096),
              (256, 256, 256), (512, 512, 512), (1024, 1024, 1024), (512, 512, 4096)]
    if quick:
        shapes = shapes[:3] + shapes[5:7]
    # Slices at least 128 deep: a slice of depth d writes and re-reads its partial C, 16 / (2d) bytes per
    # flop, which at d = 64 would ask 8 TB/s of a product running at the FP64 peak.
    cands = [(0, 1024), (4, 1024), (4, 512), (2, 512), (1, 512), (4, 256), (2, 256), (1, 256), (4, 128), (2, 128),
             (1, 128)]
    rng = np.random.default_rng(0)
    bufs = {}
    for (m, n, k) in shapes:
        bufs[(m, n, k)] = (native.DeviceArray.from_numpy(rng.uniform(-1, 1, m * k)),
                           native.DeviceArray.from_numpy(rng.uniform(-1, 1, k * n)), native.DeviceArray(m * n))
    times = {}
    for tpc, mks in cands:
        native.tune_set("splitk_tiles_per_cu", tpc)
        native.tune_set("splitk_min_kslice", mks)
        for shp, (a, b, c) in bufs.items():
            m, n, k = shp
            times[(tpc, mks, shp)] = _time_dev(
                lambda: native.dgemm_dev(0, 0, m, n, k, 1.0, a.ptr, k, b.ptr, n, 0.0, c.ptr, n, stream=native.stream(),
                                         stats=False), 3)
    best_per_shape = {shp: min(times[(t, k, shp)] for t, k in cands) for shp in bufs}
    score = {}
    for tpc, mks in cands:
        score[(tpc, mks)] = math.exp(np.mean([math.log(times[(tpc, mks, s)] / best_per_shape[s]) for s in bufs]))
    best = min(score, key=score.get)
    native.tune_set("splitk_tiles_per_cu", best[0])
    native.tune_set("splitk_min_kslice", best[1])
    return {"chosen": {"splitk_tiles_per_cu": best[0], "splitk_min_kslice": best[1]},
            "score_vs_best": {f"{t},{k}": round(v, 4) for (t, k), v in score.items()},
            "ms": {f"{t},{k} {s}": round(v, 3) for (t, k, s), v in times.items()}}


def tune_grouped(quick: bool) -> dict:
    dims, alpha, beta = batched_shapes(5000 if quick else 20000, seed=3)
    w = BatchedWorkload(dims, alpha, beta, seed=4)
    descs = w.descs()

    def timed() -> float:
        plan = native.GroupedPlan(descs)
        ms = _time_dev(lambda: plan.run(native.stream()), 5)
        plan.destroy()
        return ms

    res = {}
    for wps in (0, 1, 2, 3):
        native.tune_set("grouped_waves_per_simd", wps)
        res[wps] = timed()
    best = min(res, key=res.get)
    native.tune_set("grouped_waves_per_simd", best)
    # Tile shape per size class: problems with m (n) up to the bound run 16-row (16-column) tiles.
    shapes = {}
    for small in (0, 16, 48, 80):
        native.tune_set("grouped_small_tile", small)
        shapes[small] = timed()
    best_small = min(shapes, key=shapes.get)
    native.tune_set("grouped_small_tile", best_small)
    # Tiles of one problem dealt to one XCD together.
    chunks = {}
    for chunk in (8, 16, 64):
        native.tune_set("grouped_chunk_tiles", chunk)
        chunks[chunk] = timed()
    best_chunk = min(chunks, key=chunks.get)
    native.tune_set("grouped_chunk_tiles", best_chunk)
    return {"chosen": {"grouped_waves_per_simd": best, "grouped_small_tile": best_small,
                       "grouped_chunk_tiles": best_chunk},
            "ms": {str(k): round(v, 4) for k, v in res.items()},
            "small_tile_ms": {str(k): round(v, 4) for k, v in shapes.items()},
            "chunk_ms": {str(k): round(v, 4) for k, v in chunks.items()}}


def tune_qr_batched(quick: bool) -> dict:
    rng = np.random.default_rng(5)
    count = 5000 if quick else 20000
    ms_ = rng.integers(16, 193, count)
    mats = [rng.standard_normal(int(m) * max(1, int(m) // 2)) for m in ms_]
    flat = np.concatenate(mats)
    offs = np.concatenate([[0], np.cumsum([x.size for x in mats])[:-1]])
    toffs = np.concatenate([[0], np.cumsum([max(1, int(m) // 2) for m in ms_])[:-1]])
    da = native.DeviceArray.from_numpy(flat)
    dt = native.DeviceArray(int(toffs[-1] + max(1, int(ms_[-1]) // 2)))
    descs = [native.QrDesc(int(m), max(1, int(m) // 2), int(m), da.at(int(o)), dt.at(int(t)))
             for m, o, t in zip(ms_, offs, toffs)]
    res = {}
    for reg in (0, 32, 64, 96, 128, 160, 192):
        for threads in (512, 1024):
            for bound in (0, 2560, 4096, 5120):
                native.tune_set("qr_reg_max_m", reg)
                native.tune_set("qr_wave_max_lds", bound)
                native.tune_set("qr_block_threads", threads)
                plan = native.QrBatchedPlan(descs)
                res[(reg, bound, threads)] = _time_dev(lambda: plan.replay(native.stream()), 3)
                plan.destroy()
                da.copy_from(flat)
    best = min(res, key=res.get)
    native.tune_set("qr_reg_max_m", best[0])
    native.tune_set("qr_wave_max_lds", best[1])
    native.tune_set("qr_block_threads", best[2])
    return {"chosen": {"qr_reg_max_m": best[0], "qr_wave_max_lds": best[1], "qr_block_threads": best[2]},
            "count": count, "ms": {f"{r},{b},{t}": round(v, 3) for (r, b, t), v in res.items()}}


def tune_qr_panel(quick: bool) -> dict:
    heights = [256, 384, 512, 768, 1024, 2048, 4096] if not quick else [256, 512, 1024, 4096]
    b = 128
    rng = np.random.default_rng(9)
    res = {}
    saved = native.tune_get("qr_cqr_min_rows")
    for m in heights:
        a0 = native.DeviceArray.from_numpy(rng.standard_normal(m * b))
        work = native.DeviceArray(m * b)
        tau = native.DeviceArray(b)
        times = {}
        for name, bound in (("cholesky_qr2", 1), ("columns", 0)):
            native.tune_set("qr_cqr_min_rows", bound)

            def run():
                native.dev_copy(work.ptr, a0.ptr, m * b * 8)
                native.dgeqrf_dev(m, b, work.ptr, m, tau.ptr, b)
            times[name] = _time_dev(run, 5)
        res[m] = times
    native.tune_set("qr_cqr_min_rows", saved)
    chosen = 0
    for m in sorted(heights, reverse=True):  # smallest height from which CholeskyQR2 stays faster
        if res[m]["cholesky_qr2"] < res[m]["columns"]:
            chosen = m
        else:
            break
    native.tune_set("qr_cqr_min_rows", chosen)
    return {"chosen": {"qr_cqr_min_rows": chosen},
            "ms": {str(m): {k: round(v, 4) for k, v in t.items()} for m, t in res.items()}}


def tune_offload(cpu_gflops: float, quick: bool) -> dict:
    sizes = [32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024]
    if quick:
        sizes = sizes[::2]
    rng = np.random.default_rng(1)
    rows = []
    for n in sizes:
        a, b = rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * n)
        c = np.zeros(n * n)
        native.dgemm(0, 0, n, n, n, 1.0, a, n, b, n, 0.0, c, n)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            native.dgemm(0, 0, n, n, n, 1.0, a, n, b, n, 0.0, c, n)
        gpu_ms = (time.perf_counter() - t0) * 1e3 / reps
        cpu_ms = 2.0 * n ** 3 / (cpu_gflops * 1e9) * 1e3
        rows.append((n, 2 * n ** 3, gpu_ms, cpu_ms))
    gate = rows[-1][1]
    for i in range(len(rows)):
        if all(r[2] < r[3] for r in rows[i:]):
            gate = rows[i][1]
            break
    native.tune_set("hip_min_flops", float(gate))
    return {"chosen": {"hip_min_flops": gate}, "cpu_model_gflops": cpu_gflops,
            "rows": [{"n": n, "flops": f, "gpu_end_to_end_ms": round(g, 4), "cpu_model_ms": round(c, 4)}
                     for n, f, g, c in rows]}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(native.TUNE_TABLE))
    ap.add_argument("--cpu-gflops", type=float, default=100.0)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--only", default="", help="comma list of splitk,grouped,qr,qrpanel,offload (default all); "
                                               "the other entries keep the loaded table's values")
    args = ap.parse_args(argv)
    only = set(args.only.split(",")) if args.only else {"splitk", "grouped", "qr", "qrpanel", "offload"}
    rc, count, name = native.probe()
    if rc != native.JLC_OK:
        print(f"no usable device: {native.last_error()}", file=sys.stderr)
        return 1
    report = {"device": name}
    if "splitk" in only:
        report["splitk"] = tune_splitk(args.quick)
    if "grouped" in only:
        report["grouped"] = tune_grouped(args.quick)
    if "qr" in only:
        report["qr_batched"] = tune_qr_batched(args.quick)
    if "qrpanel" in only:
        report["qr_panel"] = tune_qr_panel(args.quick)
    if "offload" in only:
        report["offload"] = tune_offload(args.cpu_gflops, args.quick)
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    native.tune_write(str(out))
    report["table"] = out.read_text()
    print(json.dumps(report, indent=1), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
