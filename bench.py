"""Benchmark of the HIP stacked solve on one MI355X.

Headline (the one JSON line): mixed-precision refined solve of a stack of 64
systems of order 2048 with 8 right-hand sides each, cond_2(A) = 1e4:
A = U diag(s) V^T with U, V Haar orthogonal (QR of standard-normal matrices,
sign-fixed), s = logspace(0, -4, 2048); B ~ N(0, 1); PCG64 seed 3. The 64
matrices pair 8 Haar U with 8 Haar V (A_{8i+j} = U_i S V_j^T), which keeps the
host-side QR cost to 16 factorisations; every matrix has cond_2 = 1e4 exactly.

A step is one call of the refined solver on the whole stack (inputs resident
in HBM): low-precision LU, FP64 refinement to backward error <= 1e-12 for every
right-hand side, with the backward error reported per column. The same step
in plain FP64 (LU + getrs + backward-error report) is timed beside it.
`value` = systems solved per second by the faster mixed-precision mode.

The default run (--workload all) also times the other four workloads and nests
their records under "configs": plain64 (256 x 64 x 64 FP64 stacked solve,
solves/s), gemm (8192^3 FP64 DGEMM, TFLOP/s), uneven (20,000 systems of n in
[8, 256], FP64 and FP32 + refinement), serve (Poisson-arrival serving loop,
p50/p99 latency and solves/s per rate). Each record carries its own roofline
(dominant kernel class, per-launch HIP events) and CPU baseline (NumPP's
portable kernel restated, kernel-only and with NumPP's ndarray copies, one
pinned core). --workload X runs one of them (or copy: the calibration workload
for the HBM byte counters) as the single line.

--gpus above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
PROF_KINDS = {0: ("gemm_f64", "mfma", "f64"), 1: ("gemm_f32", "mfma", "f32"), 2: ("gemm_bf16split", "mfma", "bf16"),
              3: ("lu_panel", "hbm", None), 4: ("move_trsm", "hbm", None), 5: ("getrs", "hbm", None),
              6: ("residual_f64", "hbm", None), 7: ("lu_wave64", "hbm", None), 8: ("vbatch", "hbm", None)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def pin_one_core():
    try:
        cpus = sorted(os.sched_getaffinity(0))
        os.sched_setaffinity(0, {cpus[0]})
        return cpus
    except (AttributeError, OSError):
        return None


def unpin(cpus):
    if cpus:
        os.sched_setaffinity(0, set(cpus))


def cpu_timed(units, kernel_fn, numpp_fn=None):
    """Times the oracle kernel on each unit i (kernel_fn(i)) and, interleaved
    unit by unit so that both see the same cache and clock state, its
    NumPP-equivalent form with NumPP's ndarray copies (numpp_fn(i)), on one
    pinned host core. Returns (seconds, numpp_seconds)."""
    cpus = pin_one_core()
    dt, dt2 = 0.0, (0.0 if numpp_fn is not None else None)
    try:
        for i in units:
            t1 = time.perf_counter()
            kernel_fn(i)
            dt += time.perf_counter() - t1
            if numpp_fn is not None:
                t2 = time.perf_counter()
                numpp_fn(i)
                dt2 += time.perf_counter() - t2
    finally:
        unpin(cpus)
    return dt, dt2


def cpu_record(units, unit, dt, dt2, sample, numpp_units=None, extra=None):
    model, ncpu = cpu_info()
    r = {"value": units / dt, "unit": unit, "cores": 1, "kind": "port",
         "sample": f"{sample}; oracle C restatement of NumPP's portable kernel, -O3 -ffp-contract=off, "
                   f"1 of {ncpu} host cores ({model})",
         "seconds": dt}
    if dt2 is not None:
        r["numpp_equivalent_value"] = (numpp_units if numpp_units is not None else units) / dt2
        r["numpp_equivalent_note"] = ("the same kernels plus the ndarray copies NumPP makes around them "
                                      "(ndarray copy/astype element loops, to_vec/from_vec, stack)")
    if extra:
        r.update(extra)
    return r


def haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def prof_report(lib, ctypes):
    out = {}
    for kind, (name, bound, dt) in PROF_KINDS.items():
        ms, fl, by, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        lib.npph_prof_get(kind, ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by), ctypes.byref(nl))
        if nl.value:
            out[name] = {"ms": ms.value, "flops": fl.value, "bytes": by.value, "launches": nl.value,
                         "bound": bound, "dtype": dt}
    return out


def traffic_for(kind_name, mode):
    """HBM bytes per launch of a kernel class, measured by rocprofv3 counter
    runs (FETCH_SIZE + WRITE_SIZE, scaled by the factor a plain copy of known
    size gives) and recorded in profiles/traffic.json; None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    rec = table.get(f"{mode}/{kind_name}")
    return None if rec is None else {"traffic": rec["bytes_per_launch"], "traffic_source": rec["source"]}


def bench_copy(args):
    """Plain device-to-device copy of --n MiB (default 1024): the calibration
    workload for the HBM byte counters (its algorithmic traffic is exactly
    one read and one write of the buffer)."""
    import numpp_amd
    from numpp_amd import _lib
    from numpp_amd.device import DeviceArray, Timer

    lib = numpp_amd.require_device()
    mib = args.n if args.n != 2048 else 1024
    nbytes = mib << 20
    src = DeviceArray((nbytes // 8,))
    dst = DeviceArray((nbytes // 8,))
    _lib.check(lib.npph_memset(src.ptr, 1, nbytes, None))
    for _ in range(args.warmup):
        _lib.check(lib.npph_memcpy_dtod(dst.ptr, src.ptr, nbytes, None))
    lib.npph_device_sync()
    tm = Timer()
    tm.start()
    for _ in range(args.steps):
        _lib.check(lib.npph_memcpy_dtod(dst.ptr, src.ptr, nbytes, None))
    ms = tm.stop() / args.steps
    gbs = 2.0 * nbytes / (ms / 1e3) / 1e9
    return {"metric": "device copy GB/s (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": f"copy {mib} MiB device to device", "bytes": nbytes},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": 2.0 * nbytes},
            "cpu_baseline": None}


def roofline_from(prof, kind_name, extra=None):
    p = prof[kind_name]
    avg_s = p["ms"] / p["launches"] / 1e3
    if p["bound"] == "mfma":
        achieved = p["flops"] / p["launches"] / avg_s / 1e12
        peak = PEAK_TFLOPS[p["dtype"]]
        r = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
             "traffic": None}
    else:
        achieved = p["bytes"] / p["launches"] / avg_s / 1e9
        r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    r["kernel"] = kind_name
    r["avg_launch_us"] = avg_s * 1e6
    r["launches_per_step"] = p["launches"]
    if extra:
        r.update(extra)
    return r


def bench_mixed(args):
    import ctypes

    import numpp_amd
    import oracle
    from numpp_amd import _lib
    from numpp_amd import linalg as nla
    from numpp_amd.device import DeviceArray, Timer

    lib = numpp_amd.require_device()
    count, n, nrhs, cond = args.count, args.n, 8, 1e4
    rng = np.random.default_rng(3)
    t0 = time.time()
    npair = int(np.ceil(np.sqrt(count)))
    s = np.logspace(0.0, -np.log10(cond), n)
    Us = np.stack([haar(rng, n) * s for _ in range(npair)])
    Vt = np.stack([np.ascontiguousarray(haar(rng, n).T) for _ in range(npair)])
    B = rng.standard_normal((count, n, nrhs))
    log(f"inputs: {2 * npair} Haar factors of order {n} in {time.time() - t0:.1f}s")
    dUs, dVt = DeviceArray.from_host(Us), DeviceArray.from_host(Vt)
    dA = DeviceArray((count, n, n))
    for m in range(count):
        i, j = divmod(m, npair)
        _lib.check(lib.npph_dgemm_dev(n, n, n, 1.0, dUs.ptr.value + i * n * n * 8, n, 0,
                                      dVt.ptr.value + j * n * n * 8, n, 0, 0.0, dA.ptr.value + m * n * n * 8,
                                      n, 0, 1, None))
    dB = DeviceArray.from_host(B)
    dX = DeviceArray((count, n, nrhs))
    dE = DeviceArray((count, nrhs))
    dI = DeviceArray((count,), np.int32)
    dF = DeviceArray((count,), np.int32)
    del Us, Vt, dUs, dVt

    def run(prec, terms=6):
        o = nla.refine_opts(prec, tol=1e-12, max_iter=10, fallback=True, bf16_terms=terms)
        nla.gesv_refined_dev(dA, dB, dX, dE, dI, dF, o)

    results = {}
    modes = [("fp64", 0), ("fp32", 0), ("bf16", 6), ("bf16", 3)]
    wanted = set(args.modes.split(","))
    for prec, terms in modes:
        key = prec if prec != "bf16" else f"bf16x{terms}"
        if key not in wanted:
            continue
        for _ in range(args.warmup):
            run(prec, terms)
        lib.npph_device_sync()
        tm = Timer()
        lat = []
        tm.start()
        for _ in range(args.steps):
            t1 = time.perf_counter()
            run(prec, terms)
            lib.npph_device_sync()
            lat.append((time.perf_counter() - t1) * 1e3)
        ms = tm.stop() / args.steps
        x = dX.to_host()
        results[key] = {"ms": ms, "solves_per_s": count / (ms / 1e3), "berr_max": float(dE.to_host().max()),
                        "iters": dI.to_host().tolist(), "info_nonzero": int((dF.to_host() != 0).sum()),
                        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)), "x": x}
        log(f"{key}: {ms:.3f} ms/step, {results[key]['solves_per_s']:.1f} solves/s, "
            f"berr_max {results[key]['berr_max']:.2e}, iters {sorted(set(results[key]['iters']))}")
    x64 = results["fp64"]["x"] if "fp64" in results else None
    for key, r in results.items():
        r["fwd_err_vs_fp64_max"] = (None if x64 is None else float(np.max(
            np.abs(r["x"] - x64).max(axis=(1, 2)) / np.abs(x64).max(axis=(1, 2)))))
    mixed = [k for k in results if k != "fp64" and results[k]["berr_max"] <= 1e-12 and
             all(i >= 0 for i in results[k]["iters"])]
    best = min(mixed, key=lambda k: results[k]["ms"]) if mixed else min(results, key=lambda k: results[k]["ms"])

    # Profiled pass (one extra step per mode with per-kernel HIP events).
    prof = {}
    for key in dict.fromkeys([best] + (["fp64"] if "fp64" in results else [])):
        prec = "bf16" if key.startswith("bf16") else key
        terms = int(key[5:]) if key.startswith("bf16") else 0
        lib.npph_prof_reset()
        lib.npph_prof_enable(1)
        run(prec, terms)
        lib.npph_device_sync()
        lib.npph_prof_enable(0)
        prof[key] = prof_report(lib, ctypes)
    lib.npph_prof_reset()
    dom = max(prof[best], key=lambda k: prof[best][k]["ms"])
    roof = roofline_from(prof[best], dom, traffic_for(dom, best))
    step_ms = results[best]["ms"]
    roof["share_of_step"] = prof[best][dom]["ms"] / step_ms

    # CPU baseline: the oracle's FP64 LU + solve (NumPP's portable kernel,
    # restated) on one core for a bounded sample of the same workload.
    cpu = None
    if not args.no_cpu:
        k = args.cpu_sample
        A0 = dA.to_host()[:k]
        xo = [None] * k

        def kern(i):
            xo[i] = oracle.solve(A0[i], B[i])
        dt, dt2 = cpu_timed(range(k), kern, lambda i: oracle.batched_solve_numpp(A0[i:i + 1], B[i:i + 1]))
        cpu = cpu_record(k, "solves/s", dt, dt2, f"{k} of {count} systems (n={n}, 8 RHS), FP64 LU+solve, "
                         f"each also timed with NumPP's copies")
        xg = x64 if x64 is not None else results[best]["x"]
        fe = max(float(np.abs(xg[i] - xo[i]).max() / np.abs(xo[i]).max()) for i in range(k))
        results["fp64" if x64 is not None else best]["fwd_err_vs_oracle_sample"] = fe
        log(f"cpu baseline: {dt:.2f}s for {k} systems; gpu vs oracle fwd err {fe:.2e}")

    r = results[best]
    line = {
        "metric": "solves/s + p99 latency, mixed-prec refined vs FP64 LU at backward err <=1e-12",
        "value": r["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else ("f32+f64" if best == "fp32" else "bf16+f64"),
        "data": "synthetic",
        "config": {"workload": f"refined stacked solve: {count} x {n}x{n} float64, {nrhs} RHS, cond 1e4, "
                               f"Haar U/V pairs, seed 3", "count": count, "n": n, "nrhs": nrhs,
                   "mode": best, "tol": 1e-12, "parallelism": "single-gpu"},
        "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
        "fp64_solves_per_s": results["fp64"]["solves_per_s"] if x64 is not None else None,
        "fp64_ms_per_step": results["fp64"]["ms"] if x64 is not None else None,
        "speedup_vs_fp64": results["fp64"]["ms"] / r["ms"] if x64 is not None else None,
        "berr_max": r["berr_max"], "fwd_err_vs_fp64_max": r["fwd_err_vs_fp64_max"],
        "modes": {k: {kk: v for kk, v in rr.items() if kk not in ("x", "iters")} |
                  {"iters_max": max(rr["iters"]), "iters_min": min(rr["iters"])} for k, rr in results.items()},
        "roofline": roof,
        "kernel_ms_per_step": {k: {kk: round(vv["ms"], 4) for kk, vv in v.items()} for k, v in prof.items()},
        "cpu_baseline": cpu,
    }
    return line


def bench_gemm(args):
    import ctypes

    import numpp_amd
    import oracle
    from numpp_amd import _lib
    from numpp_amd.device import DeviceArray, Timer

    lib = numpp_amd.require_device()
    n = args.n if args.n != 2048 else 8192
    rng = np.random.default_rng(1)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, n))
    dA, dB, dC = DeviceArray.from_host(a), DeviceArray.from_host(b), DeviceArray((n, n))
    del a, b

    def run():
        _lib.check(lib.npph_dgemm_dev(n, n, n, 1.0, dA.ptr, n, 0, dB.ptr, n, 0, 0.0, dC.ptr, n, 0, 1, None))

    for _ in range(args.warmup):
        run()
    lib.npph_device_sync()
    tm = Timer()
    tm.start()
    for _ in range(args.steps):
        run()
    ms = tm.stop() / args.steps
    flops = 2.0 * n ** 3
    lib.npph_prof_reset()
    lib.npph_prof_enable(1)
    run()
    lib.npph_device_sync()
    lib.npph_prof_enable(0)
    prof = prof_report(lib, ctypes)
    roof = roofline_from(prof, "gemm_f64", traffic_for("gemm_f64", "gemm"))
    cpu = None
    if not args.no_cpu:
        m = 2048
        r2 = np.random.default_rng(1)
        a2, b2 = r2.standard_normal((m, m)), r2.standard_normal((m, m))
        dt, dt2 = cpu_timed(range(1), lambda i: oracle.matmul(a2, b2), lambda i: oracle.matmul_numpp(a2, b2))
        cpu = cpu_record(2.0 * m ** 3 / 1e12, "TFLOP/s", dt, dt2, "2048^3 FP64 GEMM (i-p-j loop), seed 1")
    return {"metric": "FP64 DGEMM TFLOP/s (8192^3, device-resident)", "value": flops / (ms / 1e3) / 1e12,
            "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"C = A B, {n}x{n} float64, N(0,1) seed 1", "parallelism": "single-gpu"},
            "roofline": roof, "cpu_baseline": cpu}


def bench_plain64(args):
    import ctypes

    import numpp_amd
    import oracle
    from numpp_amd import _lib
    from numpp_amd.device import DeviceArray, Timer

    lib = numpp_amd.require_device()
    rng = np.random.default_rng(0)
    a = rng.uniform(-1.0, 1.0, (256, 64, 64)) + 64.0 * np.eye(64)
    b = rng.standard_normal((256, 64))
    dA0, dB0 = DeviceArray.from_host(a), DeviceArray.from_host(b)
    dA, dB = DeviceArray((256, 64, 64)), DeviceArray((256, 64))

    def run():
        dA.copy_from_device(dA0)
        dB.copy_from_device(dB0)
        _lib.check(lib.npph_gesv_batched_dev(256, 64, 1, dA.ptr, dB.ptr, None, None, None))

    for _ in range(args.warmup):
        run()
    lib.npph_device_sync()
    tm = Timer()
    tm.start()
    for _ in range(args.steps):
        run()
    ms = tm.stop() / args.steps
    lib.npph_prof_reset()
    lib.npph_prof_enable(1)
    run()
    lib.npph_device_sync()
    lib.npph_prof_enable(0)
    prof = prof_report(lib, ctypes)
    roof = roofline_from(prof, "lu_wave64", traffic_for("lu_wave64", "plain64"))
    x = dB.to_host()
    xo, _ = oracle.batched_solve(a, b)
    reps = 20
    dt, dt2 = cpu_timed(range(reps), lambda i: oracle.batched_solve(a, b), lambda i: oracle.batched_solve_numpp(a, b))
    cpu = cpu_record(256 * reps, "solves/s", dt, dt2, f"the full 256-system batch, {reps} repetitions")
    return {"metric": "FP64 stacked solve solves/s (256 x 64x64, 1 RHS)", "value": 256 / (ms / 1e3),
            "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "256 x 64x64 float64, U[-1,1)+64I, 1 RHS, seed 0", "parallelism": "single-gpu"},
            "rel_err_vs_oracle": float(np.abs(x - xo).max() / np.abs(xo).max()),
            "roofline": roof, "cpu_baseline": cpu}

def bench_uneven(args):
    """Uneven batch: --systems independent float64 systems (default 20,000),
    n ~ integers(8, 257), nrhs ~ integers(1, 17) (both drawn as vectors first),
    then per system in order A = uniform(-1, 1, (n, n)) + n I and
    B = standard_normal((n, nrhs)); PCG64 seed 4. One planned call per step
    (size buckets, graph replay); FP64 and FP32 + FP64 refinement."""
    import ctypes

    import numpp_amd
    import oracle
    from numpp_amd.device import DeviceArray, Timer
    from numpp_amd.vbatch import VPlan

    lib = numpp_amd.require_device()
    S = args.systems
    rng = np.random.default_rng(4)
    t0 = time.time()
    ns = rng.integers(8, 257, size=S).astype(np.int32)
    nr = rng.integers(1, 17, size=S).astype(np.int32)
    offA = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)])
    offB = np.concatenate([[0], np.cumsum(ns.astype(np.int64) * nr)])
    A = np.empty(int(offA[-1]))
    B = np.empty(int(offB[-1]))
    for i in range(S):
        n = int(ns[i])
        a = rng.uniform(-1.0, 1.0, (n, n))
        a[np.diag_indices(n)] += n
        A[offA[i]:offA[i + 1]] = a.ravel()
        B[offB[i]:offB[i + 1]] = rng.standard_normal((n, int(nr[i]))).ravel()
    log(f"inputs: {S} systems, {A.nbytes / 1e9:.2f} GB of A, in {time.time() - t0:.1f}s")
    dA, dB = DeviceArray.from_host(A), DeviceArray.from_host(B)
    dX = DeviceArray(B.shape)
    dE = DeviceArray((int(nr.sum()),))
    dI = DeviceArray((S,), np.int32)
    dF = DeviceArray((S,), np.int32)
    results = {}
    plans = {}
    for key in [m for m in ("fp64", "fp32") if m in args.modes.split(",")]:
        t1 = time.time()
        plan = VPlan(ns, nr, precision=key, tol=1e-12, max_iter=10, fallback=True, autotune=True)
        plans[key] = plan
        log(f"{key}: plan with {len(plan.buckets())} buckets in {time.time() - t1:.1f}s")
        for _ in range(args.warmup):
            plan.execute(dA, dB, dX, dE, dI, dF, use_graph=True)
        lib.npph_device_sync()
        lat = []
        tm = Timer()
        tm.start()
        for _ in range(args.steps):
            t2 = time.perf_counter()
            plan.execute(dA, dB, dX, dE, dI, dF, use_graph=True)
            lib.npph_device_sync()
            lat.append((time.perf_counter() - t2) * 1e3)
        ms = tm.stop() / args.steps
        it = dI.to_host()
        results[key] = {"ms": ms, "solves_per_s": S / (ms / 1e3), "berr_max": float(dE.to_host().max()),
                        "info_nonzero": int((dF.to_host() != 0).sum()), "iters_max": int(it.max()),
                        "iters_min": int(it.min()), "p50_ms": float(np.percentile(lat, 50)),
                        "p99_ms": float(np.percentile(lat, 99)), "x": dX.to_host()}
        log(f"{key}: {ms:.3f} ms/step, {results[key]['solves_per_s']:.0f} solves/s, "
            f"berr_max {results[key]['berr_max']:.2e}, iters {it.min()}..{it.max()}")
    if "fp64" in results:
        x64 = results["fp64"]["x"]
        den = np.maximum.reduceat(np.abs(x64), offB[:-1])
        for key, r in results.items():
            num = np.maximum.reduceat(np.abs(r["x"] - x64), offB[:-1])
            r["fwd_err_vs_fp64_max"] = float((num / den).max())
    mixed = [k for k in results if k != "fp64" and results[k]["berr_max"] <= 1e-12]
    best = min(mixed, key=lambda k: results[k]["ms"]) if mixed else min(results, key=lambda k: results[k]["ms"])
    lib.npph_prof_reset()
    lib.npph_prof_enable(1)
    plans[best].execute(dA, dB, dX, dE, dI, dF, use_graph=False)
    lib.npph_device_sync()
    lib.npph_prof_enable(0)
    prof = prof_report(lib, ctypes)
    lib.npph_prof_reset()
    dom = max(prof, key=lambda k: prof[k]["ms"])
    roof = roofline_from(prof, dom, traffic_for(dom, "uneven-" + best))
    roof["share_of_step"] = prof[dom]["ms"] / results[best]["ms"]
    for p in plans.values():
        p.close()
    cpu = None
    if not args.no_cpu:
        def one(i, f):
            n, r = int(ns[i]), int(nr[i])
            return f(A[offA[i]:offA[i + 1]].reshape(1, n, n), B[offB[i]:offB[i + 1]].reshape(1, n, r))
        dt, dt2 = cpu_timed(range(S), lambda i: one(i, oracle.batched_solve),
                            lambda i: one(i, oracle.batched_solve_numpp))
        cpu = cpu_record(S, "solves/s", dt, dt2, f"all {S} systems, FP64 LU+solve, each also timed with NumPP's "
                         f"copies")
        log(f"cpu baseline: {dt:.2f}s for {S} systems")
    r = results[best]
    return {"metric": "solves/s + p99 latency, mixed-prec refined vs FP64 LU at backward err <=1e-12",
            "value": r["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": r["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if best == "fp64" else "f32+f64", "data": "synthetic",
            "config": {"workload": f"uneven batch: {S} float64 systems, n ~ U{{8..256}}, nrhs ~ U{{1..16}}, "
                                   f"A = U[-1,1) + n I, seed 4", "systems": S, "mode": best, "tol": 1e-12,
                       "parallelism": "single-gpu"},
            "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
            "fp64_solves_per_s": results["fp64"]["solves_per_s"] if "fp64" in results else None,
            "speedup_vs_fp64": results["fp64"]["ms"] / r["ms"] if "fp64" in results else None,
            "modes": {k: {kk: v for kk, v in rr.items() if kk != "x"} for k, rr in results.items()},
            "roofline": roof, "kernel_ms_per_step": {k: round(v["ms"], 4) for k, v in prof.items()},
            "cpu_baseline": cpu}


def serve_trace(R, distinct, seed=5):
    """The serving trace: R requests cycling over `distinct` generated systems
    (PCG64 `seed`): n = exp(U(ln 16, ln 4096)), nrhs = round(exp(U(0, ln 64))),
    kappa = exp(U(ln 1e2, ln 1e8)); A = U diag(s) V^T with Haar U, V and
    s = logspace(0, -log10 kappa, n) (the product on the device). With
    probability 0.3 a request (after the first of its cycle) reuses the factors
    kept by a uniformly chosen earlier request of the same cycle, with new
    right-hand sides; every other request keeps its factors under a key of its
    own. Cycling keeps the host-side QR cost of the Haar factors bounded."""
    import numpp_amd
    from numpp_amd import _lib
    from numpp_amd.device import DeviceArray

    lib = numpp_amd.require_device()
    rng = np.random.default_rng(seed)
    D = min(R, distinct)
    ns = np.exp(rng.uniform(np.log(16), np.log(4096), D)).astype(int)
    nrs = np.maximum(1, np.rint(np.exp(rng.uniform(0.0, np.log(64), D)))).astype(int)
    kap = np.exp(rng.uniform(np.log(1e2), np.log(1e8), D))
    reuse_u = rng.uniform(size=D)
    owners, mats, rhss, owner_of = [], [], [], []
    for i in range(D):
        if i > 0 and reuse_u[i] < 0.3 and owners:
            j = owners[int(rng.integers(0, len(owners)))]
            n = ns[j]
            mats.append(None)
            owner_of.append(j)
        else:
            n = int(ns[i])
            s = np.logspace(0.0, -np.log10(kap[i]), n)
            us = DeviceArray.from_host(haar(rng, n) * s)
            vt = DeviceArray.from_host(np.ascontiguousarray(haar(rng, n).T))
            da = DeviceArray((n, n))
            _lib.check(lib.npph_dgemm_dev(n, n, n, 1.0, us.ptr, n, 0, vt.ptr, n, 0, 0.0, da.ptr, n, 0, 1, None))
            mats.append(da.to_host())
            owner_of.append(i)
            owners.append(i)
        rhss.append(rng.standard_normal((n, nrs[i])))
    idx = [i % D for i in range(R)]
    cyc = [i // D for i in range(R)]
    keys = [owner_of[j] + 1 + c * D for j, c in zip(idx, cyc)]
    keeps = [0 if mats[j] is None else 1 for j in idx]
    return (np.array([rhss[j].shape[0] for j in idx]), [mats[j] for j in idx], [rhss[j] for j in idx], keys,
            keeps, [owner_of[j] + c * D for j, c in zip(idx, cyc)])


def bench_serve(args):
    """Serving loop under Poisson arrivals at each rate in --rates (requests/s),
    the same --requests requests (cycling over --distinct generated systems)
    replayed per rate (serve_trace). Budget 32e9
    bytes (32 GB = 29.8 GiB), 4 workers, batches of up to 8 same-bucket
    requests (padded orders). Mixed (FP32 + refinement) and plain FP64. The
    roofline is the dominant kernel class of one profiled solve of the trace's
    largest system (per-launch HIP events)."""
    import ctypes

    import oracle
    import numpp_amd
    from numpp_amd import _lib
    from numpp_amd.serve import Server

    lib = numpp_amd.require_device()
    R = args.requests
    t0 = time.time()
    ns, mats, rhss, keys, keeps, owner_of = serve_trace(R, args.distinct)
    nre = sum(m is None for m in mats)
    log(f"inputs: {R} requests over {min(R, args.distinct)} systems ({nre} reuse), n {ns.min()}..{ns.max()}, "
        f"in {time.time() - t0:.1f}s")
    rates = [float(x) for x in args.rates.split(",")]
    precs = [m for m in ("fp32", "fp64") if m in args.modes.split(",")] or ["fp32", "fp64"]
    runs = []

    def arrivals(lam):
        gaps = np.random.default_rng(50 + int(lam)).exponential(1.0 / lam, R)
        return np.cumsum(gaps) - gaps[0]

    for prec in precs:
        # warm-up replay: sizes pools and workspaces
        srv = Server(budget_bytes=32e9, workers=4, max_batch=8, precision=prec)
        k = min(R, 24)
        srv.replay(np.zeros(k), mats[:k], rhss[:k], keys[:k], keeps[:k])
        srv.close()
        for lam in rates:
            srv = Server(budget_bytes=32e9, workers=4, max_batch=8, precision=prec)
            out = srv.replay(arrivals(lam), mats, rhss, keys, keeps)
            st = srv.stats()
            srv.close()
            ok = out["req_status"] == 0
            berr = max(float(np.max(e)) for e in out["berr"])
            lat = out["latency_ms"][ok]
            run = {"precision": prec, "rate": lam, "requests": R, "ok": int(ok.sum()), "wall_s": out["wall_s"],
                   "solves_per_s": int(ok.sum()) / out["wall_s"], "p50_ms": float(np.percentile(lat, 50)),
                   "p99_ms": float(np.percentile(lat, 99)), "berr_max": berr,
                   "fallbacks": int((out["iters"] < 0).sum()), "batched": st["batched"],
                   "peak_gb": st["peak_bytes"] / 1e9}
            runs.append(run)
            log(f"{prec} rate {lam:g}/s: {run['solves_per_s']:.1f} solves/s, p50 {run['p50_ms']:.1f} ms, "
                f"p99 {run['p99_ms']:.1f} ms, berr_max {berr:.2e}, ok {run['ok']}/{R}, "
                f"fallbacks {run['fallbacks']}, batched {run['batched']}")
    # Kernel classes of the trace's largest system (the requests that set
    # p99), solved once with per-launch HIP events in the headline precision,
    # padded as the server pads it. Per-launch events are not recorded inside
    # a whole replay: with four worker threads a profiled replay of the full
    # trace went silent for minutes, for a cause not yet found.
    from numpp_amd import linalg as nla
    from numpp_amd.device import DeviceArray

    big = max((i for i in range(R) if mats[i] is not None), key=lambda i: mats[i].shape[0])
    nb = mats[big].shape[0]
    Nb = min(4096, -(-nb // 128) * 128)
    ap = np.eye(Nb)
    ap[:nb, :nb] = mats[big]
    bp = np.zeros((Nb, rhss[big].shape[1]))
    bp[:nb] = rhss[big]
    dA, dB = DeviceArray.from_host(ap[None]), DeviceArray.from_host(bp[None])
    dX = DeviceArray(dB.shape)
    opts = nla.refine_opts(precs[0], 1e-12, 10, True, 6)
    nla.gesv_refined_dev(dA, dB, dX, opts=opts)  # warm-up: sizes workspaces
    _lib.check(lib.npph_stream_sync(None))
    lib.npph_prof_reset()
    lib.npph_prof_enable(1)
    nla.gesv_refined_dev(dA, dB, dX, opts=opts)
    _lib.check(lib.npph_stream_sync(None))
    lib.npph_prof_enable(0)
    prof = prof_report(lib, ctypes)
    lib.npph_prof_reset()
    dom = max(prof, key=lambda k: prof[k]["ms"])
    roof = roofline_from(prof, dom, {"scope": f"one {precs[0]} refined solve of the trace's largest system, "
                                              f"n = {nb} padded to {Nb}, {rhss[big].shape[1]} RHS"})
    cpu = None
    if not args.no_cpu:
        # The same trace served offline on one core, in arrival order, until
        # about --cpu-seconds of service time: a reuse request solves with the
        # kept factors of its owner (lu_solve), every other request factors.
        svc, kept = [], {}
        cpus = pin_one_core()
        t_start = time.perf_counter()
        for i in range(R):
            if time.perf_counter() - t_start > args.cpu_seconds:
                break
            t1 = time.perf_counter()
            if mats[i] is None:
                j = owner_of[i]
                if j not in kept:
                    continue  # its owner fell outside the sample
                oracle.lu_solve(*kept[j], rhss[i])
            else:
                lu, piv, _, _, _ = oracle.lu_decompose(mats[i])
                oracle.lu_solve(lu, piv, rhss[i])
                kept[i] = (lu, piv)
            svc.append((time.perf_counter() - t1) * 1e3)
        unpin(cpus)
        cpu = cpu_record(len(svc), "solves/s", sum(svc) / 1e3, None,
                         f"the first {len(svc)} requests of the trace served one at a time in arrival order "
                         f"(about {args.cpu_seconds:g} s of service; FP64 LU+solve, reuse via lu_solve)",
                         extra={"p50_ms": float(np.percentile(svc, 50)), "p99_ms": float(np.percentile(svc, 99))})
    # headline: the highest rate the mixed-precision server sustains (completes
    # at >= 90% of the offered rate), else the lowest rate
    def sustained(prec):
        rs = [r for r in runs if r["precision"] == prec]
        good = [r for r in rs if r["solves_per_s"] >= 0.9 * r["rate"]]
        return max(good, key=lambda r: r["rate"]) if good else min(rs, key=lambda r: r["rate"])
    h = sustained(precs[0])
    plain = [r for r in runs if r["precision"] == "fp64" and r["rate"] == h["rate"]]
    return {"metric": "solves/s + p99 latency, mixed-prec refined vs FP64 LU at backward err <=1e-12",
            "value": h["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": len(runs), "warmup": 1,
            "ms_per_step": 1e3 * h["wall_s"] / R, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64" if h["precision"] == "fp32" else "f64", "data": "synthetic",
            "config": {"workload": f"serving: {R} requests per rate over {min(R, args.distinct)} systems, "
                                   f"Poisson rates {args.rates}/s, n log-U[16,4096], "
                                   f"nrhs log-U[1,64], kappa log-U[1e2,1e8], 30% reuse, budget 32e9 B, seed 5",
                       "mode": h["precision"], "rate": h["rate"], "parallelism": "single-gpu"},
            "p50_ms": h["p50_ms"], "p99_ms": h["p99_ms"],
            "fp64_solves_per_s": plain[0]["solves_per_s"] if plain else None,
            "fp64_p99_ms": plain[0]["p99_ms"] if plain else None,
            "knee": {p: sustained(p)["rate"] for p in precs},
            "runs": runs, "roofline": roof, "kernel_ms_per_replay": {k: round(v["ms"], 3) for k, v in prof.items()},
            "cpu_baseline": cpu}


def slim(rec):
    """A nested record: the fields of a bench line that describe one config."""
    keep = ("metric", "value", "unit", "ms_per_step", "steps", "warmup", "higher_is_better", "dtype", "config",
            "p50_ms", "p99_ms", "fp64_solves_per_s", "fp64_ms_per_step", "fp64_p99_ms", "speedup_vs_fp64",
            "berr_max", "modes", "knee", "runs", "rel_err_vs_oracle", "roofline", "kernel_ms_per_step",
            "kernel_ms_per_replay", "cpu_baseline", "wall_s")
    return {k: v for k, v in rec.items() if k in keep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "mixed", "gemm", "plain64", "copy", "uneven", "serve"])
    ap.add_argument("--systems", type=int, default=20000, help="uneven workload: number of systems")
    ap.add_argument("--requests", type=int, default=1000, help="serve workload: requests per rate")
    ap.add_argument("--distinct", type=int, default=240, help="serve workload: distinct generated systems")
    ap.add_argument("--rates", default="50,100,200,400", help="serve workload: Poisson rates, requests/s")
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--cpu-sample", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="serve workload: offline CPU service budget")
    ap.add_argument("--modes", default="fp64,fp32,bf16x6,bf16x3",
                    help="mixed: comma list of fp64, fp32, bf16x6, bf16x3 (uneven/serve use fp64 and fp32)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the stacked solve shards by "
              f"system across GPUs with no exchange, which this build does not launch.", file=sys.stderr)
        sys.exit(2)
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py: multi-rank launch is not supported by this single-GPU build", file=sys.stderr)
        sys.exit(2)
    fn = {"mixed": bench_mixed, "gemm": bench_gemm, "plain64": bench_plain64, "copy": bench_copy,
          "uneven": bench_uneven, "serve": bench_serve}
    if args.workload != "all":
        line = fn[args.workload](args)
    else:
        t0 = time.time()
        line = bench_mixed(args)
        log(f"mixed done at {time.time() - t0:.0f}s")
        nested = {}
        for key, name in (("1_plain64", "plain64"), ("2_gemm", "gemm"), ("4_uneven", "uneven"), ("5_serve", "serve")):
            t1 = time.time()
            nested[key] = slim(fn[name](args))
            nested[key]["wall_s"] = time.time() - t1
            log(f"{name} done at {time.time() - t0:.0f}s")
        line["configs"] = nested
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
