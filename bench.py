"""DGEMM 16384^3 benchmark on one MI355X: the FP64 MFMA kernel and the
split-precision forms (CRT form: modular INT8 products; INT8 and BF16 slice
forms), each with its moduli / slice count chosen from the inputs' exponent
spread and its error stated against a double-double oracle on sampled
entries; the roofline fraction of the value form's dominant kernel; and the
CPU baseline (a restatement of the reference's PARALLEL_MICRO path) timed on
this machine's host cores.

Workload: C = A*B, M = N = K = 16384, FP64, row-major, with
a_ij = (u - 0.5) * exp(phi * g), u ~ U[0,1), g ~ N(0,1) (same for B), phi = 0.5
by default. Inputs are generated on the host from a fixed seed, copied to HBM
once, and are resident when the timed region starts. A step is one full GEMM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--phi P] [--size N]
Prints one JSON line.

Other legs (--workload): layouts (8192^3, transposed and column-major operands, matvec), batched (100,000
uneven GEMMs in one grouped launch), qr (blocked Householder QR 32768 x 4096), qr_batched (20,000 small
QRs), wy (the compact-WY update sweeps of n = 4096 Hessenberg and bidiagonal reductions), wy_update
(single 4096 x 4096 WY updates, for counter runs) and multiply (Matrix.multiply on heap matrices).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from jlc_amd import native  # noqa: E402

FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
INT8_MFMA_PEAK_TOPS = 5000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0
METRIC = "DGEMM 16384³ TFLOP/s at stated max rel err: split-MFMA vs FP64 MFMA"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_operand(rng: np.random.Generator, rows: int, cols: int, phi: float) -> np.ndarray:
    """a_ij = (u - 0.5) * exp(phi * g), generated in row chunks to bound temporaries."""
    out = np.empty(rows * cols, dtype=np.float64)
    chunk = max(1, (1 << 24) // cols)
    for r0 in range(0, rows, chunk):
        r1 = min(rows, r0 + chunk)
        n = (r1 - r0) * cols
        u = rng.random(n)
        g = rng.standard_normal(n)
        np.multiply(u - 0.5, np.exp(phi * g), out=out[r0 * cols:r1 * cols])
    return out


def dd_sample(a: np.ndarray, b: np.ndarray, m: int, n: int, k: int, rows, cols) -> tuple:
    """Double-double dot products of C = A B on the rows x cols grid: (hi, lo, sum |a||b|)."""
    from oracle import oracle as orc
    A = a.reshape(m, k)[rows, :].copy().reshape(-1)
    B = np.ascontiguousarray(b.reshape(k, n)[:, cols]).reshape(-1)
    ii, jj = np.meshgrid(np.arange(len(rows)), np.arange(len(cols)), indexing="ij")
    return orc.dd_entries(0, 0, k, A, k, B, len(cols), ii.ravel(), jj.ravel())


def fetch_rows(dc, n: int, rows) -> np.ndarray:
    """The given rows of a row-major device array with n columns."""
    out = np.empty((len(rows), n))
    lib = native.load()
    for t, r in enumerate(rows):
        native.check(lib.jlc_memcpy_d2h(out[t].ctypes.data, dc.ptr + 8 * int(r) * n, 8 * n))
    return out


def sample_error(sample: tuple, got: np.ndarray) -> dict:
    """Max relative and componentwise error of the sampled C entries against their dd sums."""
    hi, lo, absum = sample
    exact = hi + lo
    got = got.reshape(-1)
    err = np.abs(got - hi - lo)
    nz = exact != 0.0
    rel = float(np.max(err[nz] / np.abs(exact[nz]))) if nz.any() else 0.0
    u = 2.0 ** -53
    comp = float(np.max(err / np.maximum(absum, 1e-300)) / u)
    return {"max_rel_err": rel, "max_componentwise_err_units_u": comp, "entries": int(err.size),
            "median_rel_err": float(np.median(err[nz] / np.abs(exact[nz]))) if nz.any() else 0.0}


def read_traffic(kernel_tag: str) -> float | None:
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    p = ROOT / "profiles" / f"{kernel_tag}_traffic.txt"
    if not p.exists():
        return None
    m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text(), flags=re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(max_seconds: float = 12.0) -> dict:
    """PARALLEL_MICRO restatement (oracle) on a bounded sample of the same workload:
    a 4096^3 FP64 GEMM with the same generator, median of the timed repeats."""
    from oracle import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    s = 4096
    rng = np.random.default_rng(7)
    a = make_operand(rng, s, s, 0.5)
    b = make_operand(rng, s, s, 0.5)
    c = np.zeros(s * s)
    used = orc.parallel_micro(s, s, s, 1.0, a, s, b, s, 0.0, c, s, threads)  # warm-up
    times = []
    t_end = time.perf_counter() + max_seconds
    while len(times) < 5 or (time.perf_counter() < t_end and len(times) < 20):
        t0 = time.perf_counter()
        orc.parallel_micro(s, s, s, 1.0, a, s, b, s, 0.0, c, s, threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 5 and time.perf_counter() > t_end:
            break
    med = float(np.median(times))
    rate = 2.0 * s ** 3 / med
    full = 16384
    return {"value": round(rate / 1e12, 5), "unit": "TFLOP/s", "cores": used, "kind": "port",
            "sample": f"PARALLEL_MICRO restatement (oracle/oracle.c), {s}^3 FP64 same generator, "
                      f"median of {len(times)} runs, {used} threads of {os.cpu_count()} host cores; "
                      f"at this rate the {full}^3 workload would take ~{2.0 * full ** 3 / rate:.0f} s "
                      f"(extrapolated by flops, not run)"}


def split_ran(form: dict) -> bool:
    """Whether a split form was answered by its own kernels: in automatic mode a spread the form
    cannot hold runs the FP64 MFMA kernel instead (kernel "mfma_f64", no slices)."""
    return int(form.get("slices") or 0) > 0 and form.get("kernel") != "mfma_f64"


def choose_value_form(fp64: dict, splits: list, check_error: bool = True) -> dict:
    """The form whose throughput is the bench value: the fastest split form that ran its own kernels
    and whose max relative error against the double-double oracle is at most the FP64 MFMA form's,
    else the FP64 MFMA form. A split form that fell back to the FP64 kernel is never a candidate."""
    best = fp64
    for sf in splits:
        if not split_ran(sf):
            continue
        ok = not check_error or sf["error"]["max_rel_err"] <= fp64["error"]["max_rel_err"]
        if ok and sf["tflops"] > best["tflops"]:
            best = sf
    return best


# Dispatches queued between two stream synchronisations in the one-launch-per-problem legs. Queuing
# 100,000 asynchronous launches had put 4787 packets between a queue's read and write indices under
# rocprofv3 (more than the runtime's AQL ring holds) and the queue aborted with a malformed packet;
# 256 launches in flight keep every queue far from full and cost one kernel latency per 256 launches.
PER_CALL_IN_FLIGHT = 256


def time_form(step, warmup: int, steps: int, stream: int, main: list | None = None) -> tuple[float, float]:
    """(wall ms per step, device ms per step) for steps timed launches after warmup.
    If main is a list, each step returns the call's Stats and main collects the
    dominant kernel's own duration (HIP events on the launch stream)."""
    for _ in range(warmup):
        step()
    native.synchronize()
    timer = native.Timer(stream)
    t0 = time.perf_counter()
    timer.start()
    for _ in range(steps):
        st = step()
        if main is not None:
            main.append(st.main_ms)
    dev_ms = timer.stop()
    native.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    return wall_ms / steps, dev_ms / steps


def bench_batched(args) -> int:
    """Grouped launch over `count` uneven problems vs one launch per problem."""
    from jlc_amd.workloads import BatchedWorkload, batched_shapes
    dims, alpha, beta = batched_shapes(args.count, args.seed)
    t0 = time.perf_counter()
    w = BatchedWorkload(dims, alpha, beta, seed=args.seed + 1)
    descs = w.descs()
    log(f"batched: {w.count} problems, {w.flops / 1e9:.2f} GFLOP, {w.bytes / 1e9:.2f} GB, "
        f"setup {time.perf_counter() - t0:.1f}s")
    stream = native.stream()
    t0 = time.perf_counter()
    plan = native.GroupedPlan(descs)
    plan_ms = (time.perf_counter() - t0) * 1e3
    plan.replay(stream)  # capture + first replay
    wall_ms, dev_ms = time_form(lambda: plan.replay(stream), args.warmup, args.steps, stream)
    log(f"grouped graph replay: {wall_ms:.3f} ms wall, {dev_ms:.3f} ms device; plan build {plan_ms:.1f} ms")
    # one launch per problem, same problems, through the single-GEMM entry point
    def per_call():
        for i, d in enumerate(descs):
            native.dgemm_dev(d.layout_a, d.layout_b, d.m, d.n, d.k, d.alpha, d.a, d.lda, d.b, d.ldb, d.beta, d.c,
                             d.ldc, native.MODE_FP64, stream, stats=False)
            if i % PER_CALL_IN_FLIGHT == PER_CALL_IN_FLIGHT - 1:
                native.stream_synchronize(stream)
    per_ms = None
    if not args.no_per_call:
        per_call()
        native.synchronize()
        t0 = time.perf_counter()
        per_call()
        native.synchronize()
        per_ms = (time.perf_counter() - t0) * 1e3
        log(f"per-call launches: {per_ms:.1f} ms")
    gbs = w.bytes / (dev_ms * 1e-3) / 1e9
    # The INT8 slice form of the same batch. Slice count: the library's chooser (automatic split form)
    # on a 256^3 problem from the batch's generator (uniform[-1,1]), unless --slices fixes it.
    slices = args.slices
    if not slices:
        rng = np.random.default_rng(args.seed + 2)
        a = rng.uniform(-1.0, 1.0, 256 * 256)
        b = rng.uniform(-1.0, 1.0, 256 * 256)
        slices = native.dgemm(0, 0, 256, 256, 256, 1.0, a, 256, b, 256, 0.0, np.zeros(256 * 256), 256,
                              native.MODE_SPLIT_INT8).slices
    plan_s = native.GroupedPlan(descs, native.mode_slices(native.MODE_SPLIT_INT8, slices))
    plan_s.replay(stream)
    split_wall, split_dev = time_form(lambda: plan_s.replay(stream), args.warmup, args.steps, stream)
    log(f"grouped INT8 slice form (S = {slices}): {split_wall:.3f} ms wall, {split_dev:.3f} ms device")
    split_per_ms = None
    if not args.no_per_call:
        smode = native.mode_slices(native.MODE_SPLIT_INT8, slices)

        def split_per_call():
            for i, d in enumerate(descs):
                native.dgemm_dev(d.layout_a, d.layout_b, d.m, d.n, d.k, d.alpha, d.a, d.lda, d.b, d.ldb, d.beta,
                                 d.c, d.ldc, smode, stream, stats=False)
                if i % PER_CALL_IN_FLIGHT == PER_CALL_IN_FLIGHT - 1:
                    native.stream_synchronize(stream)
        native.synchronize()
        t0 = time.perf_counter()
        split_per_call()
        native.synchronize()
        split_per_ms = (time.perf_counter() - t0) * 1e3
        log(f"per-call launches, INT8 slice form: {split_per_ms:.1f} ms")
    errors = None
    if not args.no_error:
        # Both forms from the same initial C, against double-double alpha*A*B + beta*C on every 100th problem.
        from oracle import oracle as orc
        outs = {}
        for name, pl in (("fp64", plan), ("split_int8", plan_s)):
            w.reset_c()
            native.synchronize()
            pl.run(stream)
            native.synchronize()
            outs[name] = w.dc.to_numpy()
        errs = {name: [0.0, 0.0] for name in outs}
        for i in range(0, w.count, 100):
            m, n, k = (int(x) for x in w.dims[i])
            A, B, C0, _ = w.problem(i)
            al, be = float(w.alpha[i]), float(w.beta[i])
            hi, lo = orc.dd_gemm(0, 0, m, n, k, al, A, k, B, n, be, C0, n)
            exact = hi + lo
            scale = abs(al) * (np.abs(A.reshape(m, k)) @ np.abs(B.reshape(k, n))).ravel() + abs(be) * np.abs(C0)
            for name, cc in outs.items():
                got = cc[int(w.c_off[i]):int(w.c_off[i]) + m * n]
                err = np.abs(got - hi - lo)
                nz = exact != 0.0
                if nz.any():
                    errs[name][0] = max(errs[name][0], float(np.max(err[nz] / np.abs(exact[nz]))))
                errs[name][1] = max(errs[name][1], float(np.max(err / np.maximum(scale, 1e-300))) / 2.0 ** -53)
        errors = {name: {"max_rel_err": v[0], "max_componentwise_err_units_u": v[1],
                         "problems": len(range(0, w.count, 100))} for name, v in errs.items()}
        log(f"errors vs dd (every 100th problem): {errors}")
    plan_s.destroy()
    cpu = None
    if not args.no_cpu_baseline:
        # the reference's CPU path per problem: MicroKernel restatement on one thread, one call per problem
        from oracle import oracle as orc
        done, passes = 0, 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and passes < 20:  # whole passes over the batch, ~10 s
            for i in range(w.count):
                m, n, k = (int(x) for x in w.dims[i])
                A, B, C0, _ = w.problem(i)
                c = C0.copy()
                orc.parallel_micro(m, n, k, float(w.alpha[i]), A, k, B, n, float(w.beta[i]), c, n, 1)
            done += w.count
            passes += 1
        sec = time.perf_counter() - t0
        cpu = {"value": round(done / sec, 1), "unit": "problems/s", "cores": 1, "kind": "port",
               "sample": f"{passes} pass(es) over the same {w.count} problems, PARALLEL_MICRO restatement "
                         f"(oracle/oracle.c) on 1 thread, one call per problem from Python ({sec:.1f} s)"}
        log(f"CPU restatement: {cpu['value']} problems/s")
    print(json.dumps({
        "metric": "batched uneven FP64 GEMM, grouped persistent launch (graph replay)",
        "value": round(w.count / (wall_ms * 1e-3), 1), "unit": "problems/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.count} GEMMs, M,N,K log-uniform [2,256], 10% 2/3/4 cubes, alpha,beta U[-2,2]",
                   "flops": w.flops, "bytes": w.bytes},
        "tflops": round(w.flops / (wall_ms * 1e-3) / 1e12, 3),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("grouped_f64_100k") if w.count == 100000 else None, "kernel": "grouped_f64",
                     "kernel_ms": round(dev_ms, 4)},
        "error": None if errors is None else errors["fp64"],
        "split_int8": {"slices": int(slices), "ms_per_step": round(split_wall, 4),
                       "device_ms": round(split_dev, 4), "problems_per_s": round(w.count / (split_wall * 1e-3), 1),
                       "vs_fp64_grouped": round(wall_ms / split_wall, 3),
                       "per_call_ms": None if split_per_ms is None else round(split_per_ms, 2),
                       "speedup_vs_per_call": None if split_per_ms is None else round(split_per_ms / split_wall, 1),
                       "error": None if errors is None else errors["split_int8"]},
        "per_call_ms": None if per_ms is None else round(per_ms, 2),
        "speedup_vs_per_call": None if per_ms is None else round(per_ms / wall_ms, 1),
        "plan_build_ms": round(plan_ms, 1), "cpu_baseline": cpu,
    }), flush=True)
    return 0


def bench_layouts(args) -> int:
    """FP64 MFMA across the reference's layouts and the bandwidth-bound matvec:
    (a) 8192^3 row-major; (b) gemmStridedTransA shape M=4096, N=8192, K=16384 with A
    stored K x M; (c) gemmStridedColMajorA / ColMajorB at 6000^3; matvec M=K=32768.
    Entries uniform[-1,1], alpha=1, beta=0. Each case is checked on sampled entries
    against double-double sums: componentwise error in units of u must stay <= k."""
    from oracle import oracle as orc
    rng = np.random.default_rng(args.seed)
    stream = native.stream()
    R, Cl = native.ROW, native.COL
    # (name, layout_a, layout_b, m, n, k, lda, ldb, a rows x cols as stored, b rows x cols as stored)
    cases = [("rowmajor_8192", R, R, 8192, 8192, 8192, 8192, 8192, (8192, 8192), (8192, 8192)),
             ("transA_4096x8192x16384", Cl, R, 4096, 8192, 16384, 4096, 8192, (16384, 4096), (16384, 8192)),
             ("colmajorA_6000", Cl, R, 6000, 6000, 6000, 6000, 6000, (6000, 6000), (6000, 6000)),
             ("colmajorB_6000", R, Cl, 6000, 6000, 6000, 6000, 6000, (6000, 6000), (6000, 6000))]
    results = {}
    for name, la, lb, m, n, k, lda, ldb, ash, bsh in cases:
        a = rng.uniform(-1.0, 1.0, ash[0] * ash[1])
        b = rng.uniform(-1.0, 1.0, bsh[0] * bsh[1])
        da, db, dc = native.DeviceArray.from_numpy(a), native.DeviceArray.from_numpy(b), native.DeviceArray(m * n)
        step = lambda: native.dgemm_dev(la, lb, m, n, k, 1.0, da.ptr, lda, db.ptr, ldb, 0.0, dc.ptr, n,
                                        native.MODE_FP64, stream, stats=False)
        st = native.dgemm_dev(la, lb, m, n, k, 1.0, da.ptr, lda, db.ptr, ldb, 0.0, dc.ptr, n, native.MODE_FP64,
                              stream, stats=True)
        wall_ms, dev_ms = time_form(step, args.warmup, args.steps, stream)
        tf = 2.0 * m * n * k / (dev_ms * 1e-3) / 1e12
        srng = np.random.default_rng(args.seed + len(results))
        ii = srng.integers(0, m, 256)
        jj = srng.integers(0, n, 256)
        hi, lo, absum = orc.dd_entries(la, lb, k, a, lda, b, ldb, ii, jj)
        got = dc.to_numpy().reshape(m, n)[ii, jj]
        comp = float(np.max(np.abs(got - hi - lo) / np.maximum(absum, 1e-300)) / 2.0 ** -53)
        results[name] = {"kernel": st.kernel_name, "ms_per_step": round(wall_ms, 3), "device_ms": round(dev_ms, 3),
                         "tflops": round(tf, 3), "frac_of_fp64_peak": round(tf / FP64_MFMA_PEAK_TFLOPS, 4),
                         "max_componentwise_err_units_u": round(comp, 3), "bound_units_u": k,
                         "within_bound": comp <= k}
        log(f"{name}: {dev_ms:.2f} ms device, {tf:.2f} TFLOP/s, componentwise err {comp:.2f} u (bound {k} u)")
        del da, db, dc
    # matvec y = A x, A 32768 x 32768 row-major (8 GiB), beta = 0: streams A once.
    mv = 32768
    t0 = time.perf_counter()
    A = np.empty(mv * mv)
    for r0 in range(0, mv, 4096):
        A[r0 * mv:(r0 + 4096) * mv] = rng.uniform(-1.0, 1.0, 4096 * mv)
    x = rng.uniform(-1.0, 1.0, mv)
    dA, dx, dy = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(x), native.DeviceArray(mv)
    log(f"matvec operands ready in {time.perf_counter() - t0:.1f}s")
    step = lambda: native.dgemv_dev(mv, mv, 1.0, dA.ptr, mv, dx.ptr, 1, 0.0, dy.ptr, 1, stream, stats=False)
    wall_ms, dev_ms = time_form(step, args.warmup, max(args.steps, 10), stream)
    algo_bytes = 8.0 * (mv * mv + mv + mv)
    gbs = algo_bytes / (dev_ms * 1e-3) / 1e9
    ii = np.random.default_rng(args.seed + 99).integers(0, mv, 256)
    hi, lo, absum = orc.dd_entries(native.ROW, native.ROW, mv, A, mv, x, 1, ii, np.zeros(256, dtype=np.int64))
    got = dy.to_numpy()[ii]
    comp = float(np.max(np.abs(got - hi - lo) / np.maximum(absum, 1e-300)) / 2.0 ** -53)
    results["matvec_32768"] = {"ms_per_step": round(wall_ms, 4), "device_ms": round(dev_ms, 4),
                               "algorithmic_bytes": algo_bytes, "gbs": round(gbs, 1),
                               "max_componentwise_err_units_u": round(comp, 3), "bound_units_u": mv,
                               "within_bound": comp <= mv,
                               "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                                            "traffic": read_traffic("dgemv_32768")}}
    log(f"matvec_32768: {dev_ms:.3f} ms device, {gbs:.0f} GB/s, componentwise err {comp:.2f} u")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        y = np.zeros(mv)
        reps = 0
        t0 = time.perf_counter()
        while reps < 3 or (time.perf_counter() - t0 < 8.0 and reps < 20):
            orc.matvec(mv, mv, 1.0, A, mv, x, 0.0, y)
            reps += 1
        sec = (time.perf_counter() - t0) / reps
        results["matvec_32768"]["cpu_baseline"] = {
            "value": round(algo_bytes / sec / 1e9, 2), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"the same 32768^2 matvec, SpecializedKernels.matvec restatement, 1 thread, "
                      f"mean of {reps} runs ({sec:.2f} s each)"}
        cpu = cpu_baseline()
        log(f"CPU: matvec {results['matvec_32768']['cpu_baseline']['value']} GB/s, GEMM {cpu['value']} TF")
    head = results["rowmajor_8192"]
    print(json.dumps({
        "metric": "FP64 MFMA DGEMM TFLOP/s across layouts (8192^3 row-major headline) + matvec GB/s",
        "value": head["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "layouts: 8192^3 row-major, transA 4096x8192x16384, colmajorA/B 6000^3, "
                               "matvec 32768^2; uniform[-1,1], alpha=1, beta=0"},
        "roofline": {"bound": "mfma", "achieved": head["tflops"], "peak": FP64_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": head["frac_of_fp64_peak"],
                     "traffic": read_traffic("dgemm_f64_mfma_8192"), "kernel": "dgemm_f64_mfma"},
        "cpu_baseline": cpu, "cases": results}), flush=True)
    return 0


def _graded(rng, m: int, n: int) -> np.ndarray:
    """A = U diag(s) V^T, U and V from QR of Gaussian matrices, s geometric 1 -> 1e-8."""
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (U * np.geomspace(1.0, 1e-8, n)) @ V.T


def bench_qr(args) -> int:
    """Tall blocked Householder QR: 32768 x 4096, nb = 128, FP64 panels, FP64 or split trailing updates."""
    m, n, nb = 32768, 4096, 128
    rng = np.random.default_rng(args.seed)
    stream = native.stream()
    t0 = time.perf_counter()
    # U: thin Q of a Gaussian m x n, V: Q of a Gaussian n x n, both by this library's QR.
    def q_of_gaussian(r, c):
        g = native.DeviceArray.from_numpy(rng.standard_normal((c, r)).reshape(-1))  # column-major r x c
        tau = native.DeviceArray(c)
        native.dgeqrf_dev(r, c, g.ptr, r, tau.ptr, nb)
        q = native.DeviceArray(r * c)
        native.dorgqr_dev(r, c, g.ptr, r, tau.ptr, nb, q.ptr, r)
        return q
    U = q_of_gaussian(m, n)
    V = q_of_gaussian(n, n).to_numpy().reshape(n, n)          # row j = column j of V
    Vs = (V.T * np.geomspace(1.0, 1e-8, n)).T.copy()          # columns of V scaled by s (column-major)
    dVs = native.DeviceArray.from_numpy(Vs.reshape(-1))
    A = native.DeviceArray(m * n)
    # A^T (n x m row-major = A column-major) = (V diag(s)) U^T
    native.dgemm_dev(native.COL, native.ROW, n, m, n, 1.0, dVs.ptr, n, U.ptr, m, 0.0, A.ptr, m)
    native.synchronize()
    log(f"qr: built A ({m}x{n}) in {time.perf_counter() - t0:.1f}s")
    flops = 2.0 * m * n * n - 2.0 * n ** 3 / 3.0
    work = native.DeviceArray(m * n)
    tau = native.DeviceArray(n)
    forms = {}
    update_ms = []
    for name, mode in (("fp64", native.MODE_FP64), ("split_int8", native.MODE_SPLIT_INT8),
                       ("split_crt", native.MODE_SPLIT_CRT)):
        times = []
        for it in range(args.warmup + max(1, args.steps)):
            native.dev_copy(work.ptr, A.ptr, m * n * 8, stream)
            native.synchronize()
            t1 = time.perf_counter()
            st = native.dgeqrf_dev(m, n, work.ptr, m, tau.ptr, nb, mode, stream)
            native.synchronize()
            if it >= args.warmup:
                times.append((time.perf_counter() - t1) * 1e3)
                if name == "fp64":
                    update_ms.append(st.main_ms)  # the C^T -= W2^T V^T products' device time (HIP events)
        ms = float(np.median(times))
        recon, orth = native.dqr_errors_dev(m, n, A.ptr, m, work.ptr, m, tau.ptr, nb)
        forms[name] = {"ms": round(ms, 2), "tflops": round(flops / (ms * 1e-3) / 1e12, 3),
                       "recon_err": recon, "orth_err": orth,
                       "panels": {"cholesky_qr2": st.tile_n, "column_kernels": st.tile_k}}
        log(f"qr {name}: {ms:.1f} ms, {forms[name]['tflops']} TFLOP/s, recon {recon:.3e}, orth {orth:.3e}")
    # CPU path: the oracle's Householder QR restatement on a bounded sample of the same construction.
    from oracle import oracle as orc
    cm, cn = 4096, 512
    As = _graded(np.random.default_rng(args.seed + 2), cm, cn)
    f = np.asfortranarray(As).reshape(-1, order="F").copy()
    t1 = time.perf_counter()
    ctau = orc.qr_factor(f, cm, cn)
    cpu_s = time.perf_counter() - t1
    Q = orc.qr_form_q(f, ctau, cm, cn).reshape(cn, cm).T
    R = np.triu(f.reshape(cn, cm).T[:cn])
    cpu = {"sample": f"{cm}x{cn} same construction, oracle HouseholderQR restatement, 1 thread",
           "seconds": round(cpu_s, 2), "gflops": round((2.0 * cm * cn * cn - 2.0 * cn ** 3 / 3) / cpu_s / 1e9, 3),
           "recon_err": float(np.linalg.norm(As - Q @ R) / np.linalg.norm(As)),
           "orth_err": float(np.linalg.norm(Q.T @ Q - np.eye(cn)))}
    best = min(forms.values(), key=lambda f: f["ms"])
    # Roofline of the factorization's dominant kernel, the trailing-update product C^T -= W2^T V^T: one FP64
    # MFMA GEMM launch per pair of panels [k0, k0 + 2nb) (both panels' reflectors at once), nt x mk x kc with
    # nt = n - k1 columns left, mk = m - k0 rows and kc = k1 - k0 <= 2nb: algorithmic flops 2 nt mk kc per
    # launch, timed by HIP events around each launch.
    uflops, launches = 0.0, 0
    for k0 in range(0, n, 2 * nb):
        k1 = min(k0 + 2 * nb, n)
        if n - k1 <= 0:
            continue
        uflops += 2.0 * (n - k1) * (m - k0) * (k1 - k0)
        launches += 1
    ums = float(np.mean(update_ms))
    tf = uflops / (ums * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(tf, 3), "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / FP64_MFMA_PEAK_TFLOPS, 4),
                "traffic": read_traffic("dgemm_f64_mfma_qr_update_32768x4096"),
                "kernel": "dgemm_f64_mfma (C^T -= W2^T V^T)",
                "kernel_ms": round(ums / launches, 5), "launches": launches,
                "note": f"{uflops / launches:.4g} algorithmic flops per launch (K = 2nb = 256, beta = 1: C read "
                        f"and written); the products take {ums:.1f} of the {forms['fp64']['ms']:.1f} ms (FP64 "
                        f"form); panels by CholeskyQR2 + Householder reconstruction: "
                        f"{forms['fp64']['panels']['cholesky_qr2']} of {-(-n // nb)}"}
    cpu_baseline = {"value": round(cpu["gflops"] / 1e3, 6), "unit": "TFLOP/s", "cores": 1, "kind": "port",
                    "sample": cpu["sample"] + f" ({cpu['seconds']} s)"}
    print(json.dumps({
        "metric": "blocked Householder QR 32768x4096 (nb=128), A = U diag(s) V^T, s geometric 1..1e-8",
        "value": best["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic", "config": {"workload": "qr 32768x4096 nb=128", "flops": flops},
        "forms": forms, "roofline": roofline, "cpu_baseline": cpu_baseline, "cpu_path": cpu}), flush=True)
    return 0


def bench_qr_batched(args) -> int:
    """20,000 independent QRs, m ~ U[16,192], n = m/2, one graph-replayed launch."""
    from oracle import oracle as orc
    rng = np.random.default_rng(args.seed)
    count = 20000
    ms_ = rng.integers(16, 193, count)
    mats = [None] * count
    for mm in np.unique(ms_):  # stacked numpy QRs per shape
        idx = np.nonzero(ms_ == mm)[0]
        nn = max(1, int(mm) // 2)
        U, _ = np.linalg.qr(rng.standard_normal((len(idx), int(mm), nn)))
        V, _ = np.linalg.qr(rng.standard_normal((len(idx), nn, nn)))
        A = (U * np.geomspace(1.0, 1e-8, nn)) @ np.swapaxes(V, 1, 2)
        for t, i in enumerate(idx):
            mats[i] = A[t]
    flat = np.concatenate([np.asfortranarray(M).reshape(-1, order="F") for M in mats])
    offs = np.concatenate([[0], np.cumsum([M.size for M in mats])[:-1]])
    toffs = np.concatenate([[0], np.cumsum([min(M.shape) for M in mats])[:-1]])
    da = native.DeviceArray.from_numpy(flat)
    dt = native.DeviceArray(int(toffs[-1] + min(mats[-1].shape)))
    descs = [native.QrDesc(M.shape[0], M.shape[1], M.shape[0], da.at(int(o)), dt.at(int(t)))
             for M, o, t in zip(mats, offs, toffs)]
    plan = native.QrBatchedPlan(descs)
    stream = native.stream()
    times = []
    for it in range(args.warmup + max(1, args.steps)):
        native.check(native.load().jlc_memcpy_h2d(da.ptr, flat.ctypes.data, flat.size * 8))
        native.synchronize()
        t1 = time.perf_counter()
        plan.replay(stream)
        native.synchronize()
        if it >= args.warmup:
            times.append((time.perf_counter() - t1) * 1e3)
    ms = float(np.median(times))
    got, taus = da.to_numpy(), dt.to_numpy()
    flops = float(sum(2.0 * M.shape[0] * M.shape[1] ** 2 - 2.0 * M.shape[1] ** 3 / 3 for M in mats))
    recon, orth = 0.0, 0.0
    for i, M in enumerate(mats):
        m, n = M.shape
        f = got[offs[i]:offs[i] + m * n].copy()
        Q = orc.qr_form_q(f, taus[toffs[i]:toffs[i] + n], m, n).reshape(n, m).T
        R = np.triu(f.reshape(n, m).T[:n])
        recon = max(recon, float(np.linalg.norm(M - Q @ R) / np.linalg.norm(M)))
        orth = max(orth, float(np.linalg.norm(Q.T @ Q - np.eye(n))))
    # CPU path on the first 2000 matrices
    t1 = time.perf_counter()
    for M in mats[:2000]:
        f = np.asfortranarray(M).reshape(-1, order="F").copy()
        orc.qr_factor(f, M.shape[0], M.shape[1])
    cpu_ms = (time.perf_counter() - t1) * 1e3
    print(json.dumps({
        "metric": "batched Householder QR, 20000 matrices m~U[16,192], n=m/2, one graph launch",
        "value": round(count / (ms * 1e-3), 1), "unit": "matrices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "20000 QRs, A = U diag(s) V^T per matrix", "flops": flops},
        "tflops": round(flops / (ms * 1e-3) / 1e12, 4), "max_recon_err": recon, "max_orth_err": orth,
        # the matrices sit in registers for the whole factorization: FP64 arithmetic against the FP64 peak
        "roofline": {"bound": "mfma", "achieved": round(flops / (ms * 1e-3) / 1e12, 4), "peak": FP64_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(flops / (ms * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFLOPS, 5),
                     "traffic": read_traffic("qr_batched_20000"), "kernel": "qr_batched_reg",
                     "kernel_ms": round(ms, 3),
                     "note": "each matrix in the registers of 8 threads per column, one barrier per column: "
                             "latency bound (the whole replay, every size bucket, is timed)"},
        "cpu_baseline": {"value": round(2000 / (cpu_ms * 1e-3), 1), "unit": "matrices/s", "cores": 1,
                         "kind": "port", "sample": f"first 2000 matrices, oracle HouseholderQR restatement, "
                                                   f"1 thread ({cpu_ms / 1e3:.2f} s)"},
        "cpu_path": {"sample": "first 2000 matrices, oracle HouseholderQR restatement, 1 thread",
                     "ms": round(cpu_ms, 1), "matrices_per_s": round(2000 / (cpu_ms * 1e-3), 1)}}), flush=True)
    return 0


def _wy_panel(rng: np.random.Generator, rows: int, kb: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit lower-trapezoidal V (rows x kb, row-major) with tau_j = 2 / v_j^T v_j and T from buildT's
    recurrence T[:i, i] = T[:i, :i] (-tau_i V[:, :i]^T v_i) (BlockedHessenbergQR.java:239-274)."""
    V = np.tril(0.3 * rng.standard_normal((rows, kb)), -1)
    V[np.arange(kb), np.arange(kb)] = 1.0
    G = V.T @ V
    tau = 2.0 / np.diag(G)
    T = np.zeros((kb, kb))
    for i in range(kb):
        T[i, i] = tau[i]
        if i:
            T[:i, i] = T[:i, :i] @ (-tau[i] * G[:i, i])
    return np.ascontiguousarray(V).reshape(-1), T.reshape(-1).copy()


def bench_wy(args) -> int:
    """The compact-WY trailing updates of two n = 4096 reductions, every panel in order, on device pointers:
    BlockedHessenbergQR (panel 32: H from the left, H and Q from the right, BlockedHessenbergQR.java:120-137)
    and BlockedBidiagonalization (panel 64: B[k:, k+pw:] from the left and the right,
    BlockedBidiagonalization.java:107-122, 296-316, 498-510). The panels' V and T are generated and
    resident before timing; a step is one whole sweep of updates (the panel factorizations are not timed)."""
    from oracle import oracle as orc
    n = 4096
    rng = np.random.default_rng(args.seed)
    A = rng.uniform(-1.0, 1.0, n * n)
    Q0 = np.eye(n).reshape(-1)
    sweeps = {}
    # Hessenberg: panel k, m = n - k - 1 reflector rows; H[k+1:, k:] <- Q^T H, H[:, k+1:] <- H Q, Q[:, k+1:] <- Q Q_k
    ups, vts = [], []
    for k in range(0, n - 2, 32):
        pw, m = min(32, n - 2 - k), n - k - 1
        vts.append(_wy_panel(rng, m, pw))
        p = len(vts) - 1
        ups += [("h", native.LEFT, 1, k + 1, k, m, n - k, pw, p), ("h", native.RIGHT, 0, 0, k + 1, n, m, pw, p),
                ("q", native.RIGHT, 0, 0, k + 1, n, m, pw, p)]
    sweeps["hessenberg"] = (32, ups, vts)
    # bidiagonalization: panel k, B[k:, k+pw:] from the left (V: n-k rows) and the right (V: n-k-pw rows)
    ups, vts = [], []
    for k in range(0, n, 64):
        pw = min(64, n - k)
        if k + pw >= n:
            break
        vts.append(_wy_panel(rng, n - k, pw))
        ups.append(("h", native.LEFT, 0, k, k + pw, n - k, n - k - pw, pw, len(vts) - 1))
        vts.append(_wy_panel(rng, n - k - pw, pw))
        ups.append(("h", native.RIGHT, 0, k, k + pw, n - k, n - k - pw, pw, len(vts) - 1))
    sweeps["bidiagonal"] = (64, ups, vts)

    d_a0, d_q0 = native.DeviceArray.from_numpy(A), native.DeviceArray.from_numpy(Q0)
    d_h, d_q = native.DeviceArray(n * n), native.DeviceArray(n * n)
    stream = native.stream()
    out = {}
    for name, (kb, ups, vts) in sweeps.items():
        dv = [(native.DeviceArray.from_numpy(v), native.DeviceArray.from_numpy(t)) for v, t in vts]
        flops = sum(4.0 * r * c * w + 2.0 * w * w * (c if side == native.LEFT else r)
                    for _, side, _, _, _, r, c, w, _ in ups)
        nbytes = sum(16.0 * r * c for _, _, _, _, _, r, c, _, _ in ups)

        def run(count=None):
            for tgt, side, tt, r0, c0, r, c, w, p in ups[:count]:
                base = d_h if tgt == "h" else d_q
                native.wy_apply_dev(side, tt, r, c, w, dv[p][0].ptr, w, dv[p][1].ptr, w, base.at(r0 * n + c0), n,
                                    native.MODE_FP64, stream, stats=False)

        def reset():
            native.dev_copy(d_h.ptr, d_a0.ptr, n * n * 8)
            native.dev_copy(d_q.ptr, d_q0.ptr, n * n * 8)
            native.synchronize()

        times = []
        for it in range(args.warmup + max(1, args.steps)):
            reset()
            t1 = time.perf_counter()
            run()
            native.synchronize()
            if it >= args.warmup:
                times.append((time.perf_counter() - t1) * 1e3)
        ms = float(np.median(times))
        # parity and the CPU path on the first updates of the sweep (two panels)
        sample = 6 if name == "hessenberg" else 4
        reset()
        run(sample)
        native.synchronize()
        got_h, got_q = d_h.to_numpy(), d_q.to_numpy()
        exp_h, exp_q = A.copy(), Q0.copy()
        t1 = time.perf_counter()
        for tgt, side, tt, r0, c0, r, c, w, p in ups[:sample]:
            v, t = vts[p]
            orc.wy_apply(side, tt, r, c, w, v, w, t, w, exp_h if tgt == "h" else exp_q, n, c_off=r0 * n + c0)
        cpu_ms = (time.perf_counter() - t1) * 1e3
        cpu_flops = sum(4.0 * r * c * w + 2.0 * w * w * (c if side == native.LEFT else r)
                        for _, side, _, _, _, r, c, w, _ in ups[:sample])
        err = max(float(np.max(np.abs(g - e)) / max(1e-300, float(np.max(np.abs(e)))))
                  for g, e in ((got_h, exp_h), (got_q, exp_q)))
        out[name] = {"panel": kb, "updates": len(ups), "launches": 3 * len(ups), "ms": round(ms, 3),
                     "tflops": round(flops / (ms * 1e-3) / 1e12, 3), "flops": flops,
                     "algorithmic_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "sample_updates": sample, "sample_max_rel_err": err,
                     "cpu_ms": round(cpu_ms, 1), "cpu_gflops": round(cpu_flops / (cpu_ms * 1e-3) / 1e9, 2)}
        log(f"{name}: {out[name]}")
        del dv
    hs, bd = out["hessenberg"], out["bidiagonal"]
    total_ms = hs["ms"] + bd["ms"]
    total_flops = hs["flops"] + bd["flops"]
    cpu_gflops = (hs["cpu_gflops"] + bd["cpu_gflops"]) / 2
    print(json.dumps({
        "metric": "compact-WY trailing updates of n=4096 Hessenberg (panel 32) and bidiagonal (panel 64) sweeps",
        "value": round(total_flops / (total_ms * 1e-3) / 1e12, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(total_ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "wy sweeps n=4096: Hessenberg panel 32 (3 updates per panel), bidiagonal panel 64 "
                               "(2 updates per panel)", "flops": total_flops},
        "sweeps": out,
        # Hessenberg updates (panel 32) carry kb/4 = 8 flop per algorithmic byte, below FP64's ridge of
        # 78.6 / 8 = 9.8: that sweep is priced against HBM with C read and written once per update
        "roofline": {"bound": "hbm", "achieved": hs["algorithmic_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(hs["algorithmic_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     # HBM bytes of one kb = 32 update of the whole 4096 block (its three launches), counted
                     "traffic_one_update_4096": read_traffic("wy_update_4096"),
                     "kernel": "wy_apply_dev Hessenberg sweep (wy_partial, wy_w2, wy_update per update)",
                     "kernel_ms": hs["ms"],
                     "note": "algorithmic bytes 16 * rows * cols per update (C read and written once); the "
                             "partial-W kernel reads C once more, so 24 B per element reach HBM and 2/3 of the "
                             "peak is this form's ceiling"},
        "cpu_baseline": {"value": round(cpu_gflops / 1e3, 4), "unit": "TFLOP/s", "cores": orc.threads(),
                         "kind": "port", "sample": "first 2 panels of each sweep (10 updates), oracle "
                                                   f"applyBlockLeft/Right restatement, {orc.threads()} threads"}}),
        flush=True)
    return 0


def bench_wy_update(args) -> int:
    """Single compact-WY updates of one 4096 x 4096 device block (row stride 4096): kb = 32 and 64, left then
    right, three calls each (the counter runs of profiles/wy_update_4096_traffic.txt use this leg)."""
    n = 4096
    rng = np.random.default_rng(args.seed)
    a = native.DeviceArray.from_numpy(rng.uniform(-1.0, 1.0, n * n))
    out = {}
    for kb in (32, 64):
        v, t = _wy_panel(rng, n, kb)
        dv, dt = native.DeviceArray.from_numpy(v), native.DeviceArray.from_numpy(t)
        for side, name in ((native.LEFT, "left"), (native.RIGHT, "right")):
            ms = [native.wy_apply_dev(side, 0, n, n, kb, dv.ptr, kb, dt.ptr, kb, a.ptr, n).kernel_ms
                  for _ in range(3)]
            out[f"{name}_kb{kb}"] = {"ms": [round(x, 4) for x in ms], "gbs_algorithmic": round(16.0 * n * n / (min(ms) * 1e-3) / 1e9, 1)}
            log(f"wy update {name} kb {kb}: {out[f'{name}_kb{kb}']}")
    print(json.dumps({"metric": "compact-WY update of a 4096 x 4096 block, ms per call (three calls per case)",
                      "unit": "ms", "dtype": "f64", "data": "synthetic", "cases": out}), flush=True)
    return 0


def bench_multiply(args) -> int:
    """Matrix.multiply on heap matrices (the reference's facade, Matrix.java:1120-1122, routed by
    BlockedMultiply.multiply to the HIP plug): a new row-major C = A B, A and B staged over PCIe and C copied
    back inside every call. n = 4096 and 8192, U[-1, 1) entries; median wall clock of the timed calls, and
    the device kernel time of one call through jlc_dgemm's stats beside it."""
    from jlc_amd.compute import DispatchPolicy, Matrix
    rng = np.random.default_rng(args.seed)
    out = {}
    for n in (4096, 8192):
        a = Matrix(n, n, rng.uniform(-1.0, 1.0, n * n))
        b = Matrix(n, n, rng.uniform(-1.0, 1.0, n * n))
        assert DispatchPolicy.defaultPolicy().selectForMultiply(n, n, n).value == "HIP"
        times = []
        for it in range(args.warmup + max(1, args.steps)):
            t1 = time.perf_counter()
            c = a.multiply(b)
            if it >= args.warmup:
                times.append((time.perf_counter() - t1) * 1e3)
        ms = float(np.median(times))
        st = native.dgemm(native.ROW, native.ROW, n, n, n, 1.0, a.getRawData(), n, b.getRawData(), n, 0.0,
                          np.empty(n * n), n)
        # spot check: 64 entries of C against numpy's product of those rows and columns
        idx = rng.integers(0, n, (64, 2))
        ref = np.einsum("ij,ji->i", a.to_numpy()[idx[:, 0]], b.to_numpy()[:, idx[:, 1]])
        got = c.to_numpy()[idx[:, 0], idx[:, 1]]
        out[str(n)] = {"ms": round(ms, 3), "tflops_incl_pcie": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 3),
                       "kernel_ms": round(st.kernel_ms, 3), "call_ms": round(st.total_ms, 3),
                       "max_rel_err_64_entries": float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))}
        log(f"Matrix.multiply {n}: {out[str(n)]}")
    big = out["8192"]
    print(json.dumps({
        "metric": "Matrix.multiply 8192x8192 (heap matrices, PCIe staging inside the call)",
        "value": big["tflops_incl_pcie"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": big["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": {"workload": "Matrix.multiply n = 4096 and 8192 through BlockedMultiply"},
        "sizes": out}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--phi", type=float, default=0.5)
    ap.add_argument("--phi-sweep", default="0.5,2,8",
                    help="exponent spreads reported under phi_sweep (every form, 3 timed steps each); '' for none")
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=20260603)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="time only the FP64 MFMA form")
    ap.add_argument("--slices", type=int, default=0, help="INT8 slice count for the split form (0 = auto)")
    ap.add_argument("--moduli", type=int, default=0, help="moduli count for the CRT form (0 = auto)")
    ap.add_argument("--forms", default="", help="split forms to time, comma list of crt,int8,bf16 (default all)")
    ap.add_argument("--host-path", action="store_true",
                    help="also time the host-pointer entry point (PCIe staging included) on the value form")
    ap.add_argument("--workload", choices=["dgemm", "layouts", "batched", "qr", "qr_batched", "wy", "wy_update", "multiply"], default="dgemm",
                    help="dgemm: the 16384^3 metric workload; batched: the grouped-GEMM configuration")
    ap.add_argument("--count", type=int, default=100000, help="problems in the batched workload")
    ap.add_argument("--no-per-call", action="store_true",
                    help="batched workload: skip the one-launch-per-problem comparison (counter runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel over MIB MiB three times (byte-counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2
    rc, count, name = native.probe()
    if rc != native.JLC_OK:
        print(f"no usable device: {native.last_error()}", file=sys.stderr)
        return 1
    log(f"device {name}")
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = native.DeviceArray(nbytes // 8), native.DeviceArray(nbytes // 8)
        for _ in range(3):
            native.dev_copy(dst.ptr, src.ptr, nbytes)
        native.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 3}), flush=True)
        return 0

    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "layouts":
        return bench_layouts(args)
    if args.workload == "qr":
        return bench_qr(args)
    if args.workload == "qr_batched":
        return bench_qr_batched(args)
    if args.workload == "wy":
        return bench_wy(args)
    if args.workload == "wy_update":
        return bench_wy_update(args)
    if args.workload == "multiply":
        return bench_multiply(args)

    m = n = k = args.size
    da, db, dc = native.DeviceArray(m * k), native.DeviceArray(k * n), native.DeviceArray(m * n)
    stream = native.stream()
    flops = 2.0 * m * n * k
    wanted = set(args.forms.split(",")) if args.forms else {"crt", "int8", "bf16"}
    split_modes = [] if args.no_split else [
        (name, mode) for key, name, mode in (
            ("crt", "split_crt", native.mode_slices(native.MODE_SPLIT_CRT, args.moduli)),
            ("int8", "split_int8", native.mode_slices(native.MODE_SPLIT_INT8, args.slices)),
            ("bf16", "split_bf16", native.mode_slices(native.MODE_SPLIT_BF16, 0))) if key in wanted]

    def measure_phi(phi: float, steps: int, warmup: int) -> tuple[dict, list]:
        """Every form on the workload with spread phi: (FP64 form, split forms), each with its dd error."""
        rng = np.random.default_rng(args.seed)
        t0 = time.perf_counter()
        a = make_operand(rng, m, k, phi)
        b = make_operand(rng, k, n, phi)
        da.copy_from(a)
        db.copy_from(b)
        srng = np.random.default_rng(args.seed + 1)
        rows = np.sort(srng.choice(m, 64, replace=False))
        cols = np.sort(srng.choice(n, 64, replace=False))
        sample = None if args.no_error else dd_sample(a, b, m, n, k, rows, cols)
        del a, b
        log(f"phi={phi}: inputs generated and resident in {time.perf_counter() - t0:.1f}s")

        def run_form(name: str, mode: int) -> dict:
            def step():
                return native.dgemm_dev(native.ROW, native.ROW, m, n, k, 1.0, da.ptr, k, db.ptr, n, 0.0, dc.ptr, n,
                                        mode, stream, stats=True)
            st = step()
            main = []
            wall_ms, dev_ms = time_form(step, warmup, steps, stream, main)
            form = {"form": name, "kernel": st.kernel_name, "slices": st.slices, "ms_per_step": round(wall_ms, 3),
                    "device_ms": round(dev_ms, 3), "main_kernel_ms": round(float(np.mean(main)), 3),
                    "tflops": round(flops / (wall_ms * 1e-3) / 1e12, 3)}
            if st.kernel_name == "split_crt":
                form["bits"] = [st.bits_a, st.bits_b]
            if name != "fp64_mfma" and st.kernel_name != "mfma_f64":
                crt = name == "split_crt"
                pairs = st.slices if crt else st.slices * (st.slices + 1) // 2
                ops = pairs * flops
                peak = BF16_MFMA_PEAK_TFLOPS if name == "split_bf16" else INT8_MFMA_PEAK_TOPS
                unit = "TFLOP/s" if name == "split_bf16" else "TOP/s"
                what = (f"{pairs} modular INT8 products of 2MNK ops (bits {st.bits_a}/{st.bits_b}); time includes "
                        f"residues and CRT reconstruction") if crt else \
                    f"{pairs} slice products of 2MNK ops; time includes slicing"
                kms = form["main_kernel_ms"]
                form["roofline"] = {"bound": "mfma", "achieved": round(ops / (kms * 1e-3) / 1e12, 1),
                                    "peak": peak, "unit": unit, "frac": round(ops / (kms * 1e-3) / 1e12 / peak, 4),
                                    "kernel_ms": kms, "note": what + "; achieved over the product kernel alone"}
            log(f"phi={phi} {name}: {wall_ms:.2f} ms/step wall, {dev_ms:.2f} ms device, {form['tflops']:.2f} "
                f"TFLOP/s, kernel {st.kernel_name}{' slices=%d' % st.slices if st.slices else ''}")
            if sample is not None:
                form["error"] = sample_error(sample, fetch_rows(dc, n, rows)[:, cols])
                log(f"phi={phi} {name} error vs dd oracle: {form['error']}")
            return form

        fp64 = run_form("fp64_mfma", native.MODE_FP64)
        splits = [run_form(name, mode) for name, mode in split_modes]
        for sf in splits:
            if not split_ran(sf):
                sf["fallback"] = "fp64"
        return fp64, splits

    fp64, splits = measure_phi(args.phi, args.steps, args.warmup)
    # value: best throughput among split forms that ran their own kernels and whose max relative error is
    # at most the FP64 MFMA form's; forms that fell back to the FP64 kernel are labelled as such.
    best = choose_value_form(fp64, splits, check_error=not args.no_error)
    # roofline of the value form's dominant kernel, timed alone with HIP events in the timed loop
    if best is fp64:
        achieved = flops / (fp64["main_kernel_ms"] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4),
                    "traffic": read_traffic("dgemm_f64_mfma_16384"), "kernel": "dgemm_f64_mfma",
                    "kernel_ms": fp64["main_kernel_ms"]}
    else:
        kname = {"split_crt": "crt_gemm_i8", "split_int8": "split_gemm_i8", "split_bf16": "split_gemm_bf16"}
        roofline = dict(best["roofline"])
        roofline["kernel"] = kname[best["form"]]
        roofline["traffic"] = read_traffic(f"{kname[best['form']]}_16384")
    out = {
        "metric": METRIC,
        "value": best["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"dgemm {m}x{n}x{k} row-major, a=(u-0.5)*exp(phi*g), phi={args.phi}",
                   "M": m, "N": n, "K": k, "alpha": 1.0, "beta": 0.0, "phi": args.phi, "form": best["form"],
                   "value_rule": "best TFLOP/s among forms with max rel err <= the FP64 MFMA form's",
                   "parallelism": "single_gpu"},
        "error": best.get("error"),
        "forms": {f["form"]: f for f in [fp64] + splits},
        "roofline": roofline,
        "speedup_vs_fp64": {f["form"]: round(f["tflops"] / fp64["tflops"], 3) for f in splits if split_ran(f)},
    }
    sweep = [float(x) for x in args.phi_sweep.split(",") if x.strip()] if args.phi_sweep else []
    if sweep:
        out["phi_sweep"] = {}
        for phi in sweep:
            if phi == args.phi:
                f64, spl = fp64, splits
            else:
                f64, spl = measure_phi(phi, min(args.steps, 3), 1)
            entry = {"fp64_mfma": f64, "value_form": choose_value_form(f64, spl, not args.no_error)["form"]}
            for sf in spl:
                sf = dict(sf)
                if split_ran(sf):
                    sf["speedup_vs_fp64"] = round(sf["tflops"] / f64["tflops"], 3)
                entry[sf["form"]] = sf
            out["phi_sweep"][str(phi)] = entry
    if args.host_path:
        mode = {"fp64_mfma": native.MODE_FP64, "split_crt": native.MODE_SPLIT_CRT,
                "split_int8": native.MODE_SPLIT_INT8, "split_bf16": native.MODE_SPLIT_BF16}[best["form"]]
        c = np.zeros(m * n)
        native.dgemm(native.ROW, native.ROW, m, n, k, 1.0, a, k, b, n, 0.0, c, n, mode)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            native.dgemm(native.ROW, native.ROW, m, n, k, 1.0, a, k, b, n, 0.0, c, n, mode)
        hms = (time.perf_counter() - t0) * 1e3 / reps
        out["host_path"] = {"form": best["form"], "ms_per_call": round(hms, 2),
                            "tflops": round(flops / (hms * 1e-3) / 1e12, 3),
                            "note": "jlc_dgemm on host arrays: staging of A, B (and C) over PCIe included"}
        log(f"host-pointer path: {hms:.1f} ms, {out['host_path']['tflops']} TFLOP/s")
    if not args.no_cpu_baseline:
        log("timing CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
