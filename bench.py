#!/usr/bin/env python3
"""Benchmark: FP64 QR::Factor (blocked Householder QR) of an 8192 x 8192
matrix on one MI355X through the rnpx C ABI.

A step is one QR::Factor of the 8192 x 8192 N(0,1) matrix (RNP Normal01
stream, iseed {1,2,3,5}), already resident in HBM: each step copies the
pristine matrix into the work buffer on the device and factors it.  The copy
is inside the timed region (so the figure is slightly conservative).

value   = algorithmic flops (4mnk - 2(m+n)k^2 + 4k^3/3) / ms_per_step
roofline: the trailing-update GEMM (C -= V Z, FP64 MFMA), timed live with
          HIP events on the library's stream over the timed steps (the only
          kernel carrying events there; the per-family breakdown comes from
          two further, untimed steps with every family instrumented).
cpu_baseline: the reference's own QR::Factor (built from its sources against
          the image's MKL into oracle/_ref) on the host cores, the whole
          8192 x 8192 matrix, all cores best of 3, plus a one-thread best of 3
          on a 2048^2 sample; the oracle's C restatement where that library
          is absent.
check:    ||A - QR|| / (n eps ||A||) and ||Q^T Q - I|| / (n eps) of the timed
          factorization, on the device, outside the timed region.

Extra legs (in "extra", same line): complex<double> 4096^2 and tall
32768 x 4096 Factor; the batched form (10,000 least-squares problems, m
log-uniform in [16,256], n uniform in [8,m]); mixed-precision vs plain FP64
least squares on 16384 x 2048 with s log-spaced 1..1e-6; a Poisson stream of
least-squares requests through the server with p50/p99 latency.

Prints one JSON line.  Only one GPU: --gpus N with N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix peak (published)
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E (published)


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "qr_factor_8192_kernels.md")


def update_traffic_bytes():
    """HBM bytes per trailing-update launch from the committed FETCH_SIZE /
    WRITE_SIZE counter summary (8192^2 workload), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def qr_flops(m, n, cplx=False):
    k = min(m, n)
    f = 4.0 * m * n * k - 2.0 * (m + n) * k * k + 4.0 * k ** 3 / 3.0
    return 4.0 * f if cplx else f


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m, n, seed, budget_s=15.0):
    """The reference's own QR::Factor (RNP/LA/QR.hpp:129-197, built from its
    sources against the image's MKL into oracle/_ref) on the host cores:
    the whole m x n matrix of the GPU step (the same RNP Normal01 stream),
    all cores, best of 3; and one thread, best of 3, on a bounded 2048^2
    sample.  Without that library: the oracle (a C restatement of the same
    algorithm) over the first panels of the same matrix."""
    import oracle as O
    from oracle import refqr
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    a0 = O.generate_matrix(O.NORMAL01, m, n, np.array(seed, dtype=np.int32))
    if refqr.available():
        def best_of(a_src, reps):
            best = float("inf")
            tau = np.zeros(min(a_src.shape), dtype=np.float64)
            for _ in range(reps):
                a = np.asfortranarray(a_src.copy())
                t0 = time.perf_counter()
                refqr.factor_inplace(a, tau)
                best = min(best, time.perf_counter() - t0)
            return best
        refqr.set_threads(threads)
        t_all = best_of(a0, 3)
        s = 2048
        a1 = np.asfortranarray(a0[:s, :s])
        refqr.set_threads(1)
        t_one = best_of(a1, 3)
        refqr.set_threads(threads)
        fl, fl1 = qr_flops(m, n), qr_flops(s, s)
        return {"value": round(fl / t_all / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "reference",
                "sample": f"RNP QR::Factor (nb=32, nx=128) of the whole {m}x{n} matrix, {threads} threads, "
                          f"best of 3: {t_all:.2f} s",
                "one_thread": {"value": round(fl1 / t_one / 1e9, 3), "unit": "GFLOP/s",
                               "sample": f"RNP QR::Factor of the leading {s}x{s} block, 1 thread, best of 3: "
                                         f"{t_one:.3f} s"},
                "blas": refqr.blas_name(), "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    O.set_threads(threads)
    nb = 32
    a = np.asfortranarray(a0.copy())
    t0 = time.perf_counter()
    O.factor_panels(a, nb, 1)
    t1 = time.perf_counter() - t0
    P = int(max(2, min(64, budget_s / max(t1, 1e-3))))
    a = np.asfortranarray(a0.copy())
    t0 = time.perf_counter()
    O.factor_panels(a, nb, P)
    dt = time.perf_counter() - t0
    K = P * nb
    fl = qr_flops(m, n) - qr_flops(m - K, n - K)
    return {"value": round(fl / dt / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of RNP QR::Factor, nb=32, OpenMP gemm) on the first {P} "
                      f"panels ({K} columns) of the same {m}x{n} matrix: {fl / 1e9:.1f} GFLOP in {dt:.2f} s "
                      f"(reference library unavailable: {refqr.why_unavailable()})",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count()}


def _ev_time(fn, reps, stream_sync):
    """Mean wall time of fn() over reps runs (each ended by a device sync)."""
    fn()
    stream_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    stream_sync()
    return (time.perf_counter() - t0) / reps


def leg_factor(R, torch, m, n, cplx, reps=3):
    h = R.Handle()
    dt = torch.complex128 if cplx else torch.float64
    a0 = R.colmajor_empty(m, n, dt)
    h.larnv(2, [3, 1, 4, 1], a0)
    a = R.colmajor_empty(m, n, dt)
    tau = torch.empty(min(m, n), dtype=dt, device="cuda")
    h.synchronize()
    stream = torch.cuda.ExternalStream(h.stream)

    def run():
        with torch.cuda.stream(stream):
            a.copy_(a0)
        h.geqrf(a, tau)

    t = _ev_time(run, reps, h.synchronize)
    fl = qr_flops(m, n, cplx)
    return {"shape": [m, n], "dtype": "c128" if cplx else "f64", "ms": round(t * 1e3, 3),
            "gflops": round(fl / t / 1e9, 1)}


def leg_lq(R, torch, m=4096, n=8192, reps=3):
    """LQ::Factor of a wide m x n matrix (RNP/LA/LQ.hpp:126), beside QR::Factor
    of its n x m transpose: the difference is the two conjugate transposes."""
    h = R.Handle()
    a0 = R.colmajor_empty(m, n, torch.float64)
    h.larnv(2, [2, 7, 1, 8], a0)
    a = R.colmajor_empty(m, n, torch.float64)
    at = R.colmajor_empty(n, m, torch.float64)
    tau = torch.empty(min(m, n), dtype=torch.float64, device="cuda")
    h.synchronize()
    stream = torch.cuda.ExternalStream(h.stream)

    def run_lq():
        with torch.cuda.stream(stream):
            a.copy_(a0)
        h.gelqf(a, tau)

    def run_qr():
        with torch.cuda.stream(stream):
            at.copy_(a0.t())
        h.geqrf(at, tau)

    t_lq = _ev_time(run_lq, reps, h.synchronize)
    t_qr = _ev_time(run_qr, reps, h.synchronize)
    fl = qr_flops(n, m)
    return {"shape": [m, n], "dtype": "f64", "lq_ms": round(t_lq * 1e3, 3), "lq_gflops": round(fl / t_lq / 1e9, 1),
            "qr_of_transpose_ms": round(t_qr * 1e3, 3)}


def leg_qlrq(R, torch, m=8192, n=4096, reps=3):
    """QL::Factor (RNP/LA/QL.hpp:126) of a tall m x n matrix and RQ::Factor
    (RNP/LA/RQ.hpp:131) of its wide transpose, beside QR::Factor of the same
    tall matrix: the difference is the row and column reversals."""
    h = R.Handle()
    a0 = R.colmajor_empty(m, n, torch.float64)
    h.larnv(2, [3, 1, 4, 1], a0)
    a = R.colmajor_empty(m, n, torch.float64)
    at = R.colmajor_empty(n, m, torch.float64)
    tau = torch.empty(min(m, n), dtype=torch.float64, device="cuda")
    h.synchronize()
    stream = torch.cuda.ExternalStream(h.stream)

    def run(fn, dst, src):
        def f():
            with torch.cuda.stream(stream):
                dst.copy_(src)
            fn(dst, tau)
        return f

    t_ql = _ev_time(run(h.geqlf, a, a0), reps, h.synchronize)
    t_rq = _ev_time(run(h.gerqf, at, a0.t()), reps, h.synchronize)
    t_qr = _ev_time(run(h.geqrf, a, a0), reps, h.synchronize)
    fl = qr_flops(m, n)
    return {"shape": [m, n], "dtype": "f64", "ql_ms": round(t_ql * 1e3, 3), "ql_gflops": round(fl / t_ql / 1e9, 1),
            "rq_of_transpose_ms": round(t_rq * 1e3, 3), "qr_ms": round(t_qr * 1e3, 3)}


def leg_lu(R, torch, n=8192, reps=3):
    """LU::Factor (RNP/LA/LU.hpp:77) of an n x n N(0,1) matrix through the C
    ABI with the matrix resident on the device (pivots returned to the host),
    and LU::Solve with 16 right-hand sides; flops 2n^3/3 and 2 n^2 nrhs."""
    import ctypes
    h = R.Handle()
    a0 = R.colmajor_empty(n, n, torch.float64)
    h.larnv(2, [1, 6, 1, 8], a0)
    a = R.colmajor_empty(n, n, torch.float64)
    h.synchronize()
    L = R.lib()
    piv = np.zeros(n, dtype=np.uint64)
    pp = piv.ctypes.data_as(ctypes.c_void_p)
    ts = []
    for r in range(reps + 1):
        a.copy_(a0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = L.rnpx_dgetrf(n, n, ctypes.c_void_p(a.data_ptr()), n, pp)
        torch.cuda.synchronize()
        if r:
            ts.append(time.perf_counter() - t0)
    assert info == 0, info
    b = torch.randn(n, 16, dtype=torch.float64, device="cuda").t().contiguous().t()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = L.rnpx_dgetrs(b"N", n, 16, ctypes.c_void_p(a.data_ptr()), n, pp, ctypes.c_void_p(b.data_ptr()), n)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    assert st == 0, st
    t = float(np.median(ts))
    return {"shape": [n, n], "dtype": "f64", "factor_ms": round(t * 1e3, 3),
            "factor_gflops": round((2.0 * n ** 3 / 3.0) / t / 1e9, 1), "solve16_ms": round(t_solve * 1e3, 3)}


def c3_shapes(count, seed):
    rng = np.random.default_rng(seed)
    m = np.rint(np.exp(rng.uniform(np.log(16), np.log(256), count))).astype(int)
    n = np.array([rng.integers(8, mm + 1) for mm in m])
    return list(zip(m.tolist(), n.tolist()))


def leg_batched(R, torch, cplx, count=10000, reps=5):
    shapes = c3_shapes(count, 0)
    desc, na, nt, nbb = R.make_batch(shapes, nrhs=1)
    dt = torch.complex128 if cplx else torch.float64
    h = R.Handle()
    A0 = torch.empty(na, dtype=dt, device="cuda")
    B0 = torch.empty(nbb, dtype=dt, device="cuda")
    h.larnv(2, [7, 7, 7, 7], A0)
    h.larnv(2, [9, 9, 9, 9], B0)
    A, B = A0.clone(), B0.clone()
    tau = torch.empty(nt, dtype=dt, device="cuda")
    h.synchronize()
    plan = h.batched_plan(desc, A, tau, B)
    stream = torch.cuda.ExternalStream(h.stream)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for r in range(reps + 1):
        with torch.cuda.stream(stream):
            A.copy_(A0)
            B.copy_(B0)
            start.record(stream)
        plan.run()
        with torch.cuda.stream(stream):
            end.record(stream)
        h.synchronize()
        if r:
            times.append(start.elapsed_time(end) * 1e-3)
    plan.close()
    t = float(np.median(times))
    fl = sum(qr_flops(m, nn, cplx) + (4 if cplx else 1) * (4.0 * m * nn - 2.0 * nn * nn + nn * nn) for m, nn in shapes)
    by = (16 if cplx else 8) * sum(2 * m * nn + 2 * m for m, nn in shapes)
    return {"problems": count, "dtype": "c128" if cplx else "f64", "ms": round(t * 1e3, 3),
            "solves_per_s": round(count / t, 1), "gflops": round(fl / t / 1e9, 1),
            "hbm_gbs": round(by / t / 1e9, 1), "fp64_frac": round(fl / t / 1e12 / FP64_MFMA_PEAK_TFLOPS, 4)}


def conditioned_dev(R, torch, h, m, n, cond, seed):
    """A = U diag(s) V^T on the device: U, V the Q factors (this library's QR)
    of Gaussian matrices, s log-spaced from 1 to 1/cond."""
    def q_of(rows, cols, sd):
        g = R.colmajor_empty(rows, cols, torch.float64)
        h.larnv(2, [sd % 4096, 5, 7, 11], g)
        tau = torch.empty(cols, dtype=torch.float64, device="cuda")
        h.geqrf(g, tau)
        h.orgqr(g, tau)
        return g
    u = q_of(m, n, seed)
    v = q_of(n, n, seed + 1)
    h.synchronize()
    s = torch.logspace(0, -np.log10(cond), n, dtype=torch.float64, device="cuda")
    a = R.colmajor_empty(m, n, torch.float64)
    a.copy_((u * s) @ v.t())
    torch.cuda.synchronize()
    return a


def leg_mixed_ls(R, torch, m=16384, n=2048, cond=1e6, reps=3):
    h = R.Handle()
    a = conditioned_dev(R, torch, h, m, n, cond, 21)
    xt = torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    b = R.colmajor_empty(m, 1, torch.float64)
    b.copy_(a @ xt + 1e-8 * torch.randn(m, 1, dtype=torch.float64, device="cuda",
                                        generator=torch.Generator("cuda").manual_seed(2)))
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(h.stream)
    af = R.colmajor_empty(m, n, torch.float64)
    bf = R.colmajor_empty(m, 1, torch.float64)
    tau = torch.empty(n, dtype=torch.float64, device="cuda")

    def plain():
        with torch.cuda.stream(stream):
            af.copy_(a)
            bf.copy_(b)
        h.gels(af, tau, bf)

    t_plain = _ev_time(plain, reps, h.synchronize)
    x64 = bf[:n].clone()
    be64 = h.ls_berr(a, b, x64)
    x = R.colmajor_empty(n, 1, torch.float64)
    info = {}

    def mixed():
        info["r"] = h.gels_mixed(a, b, x, max_iter=30, tol=1e-13, allow_notconv=True)

    t_mixed = _ev_time(mixed, reps, h.synchronize)
    st, iters, bem = info["r"]
    fwd = float((x - x64).abs().max() / x64.abs().max())
    return {"shape": [m, n], "cond": cond,
            "plain_fp64": {"ms": round(t_plain * 1e3, 3), "solves_per_s": round(1.0 / t_plain, 2),
                           "berr": be64},
            "mixed": {"ms": round(t_mixed * 1e3, 3), "solves_per_s": round(1.0 / t_mixed, 2), "iters": iters,
                      "berr": bem, "status": st, "rel_diff_vs_fp64": fwd},
            "speedup": round(t_plain / t_mixed, 3)}


def leg_serving(R, torch, count=120, rate=40.0, seed=5, mode=1):
    """Poisson arrivals at `rate` requests/s of least-squares requests: m
    log-uniform in [64, 8192], n uniform in [m/8, m/2], 1-16 right-hand
    sides, condition number log-uniform in [1e2, 1e8] (s log-spaced)."""
    rng = np.random.default_rng(seed)
    h = R.Handle()
    reqs = []
    for i in range(count):
        m = int(np.rint(np.exp(rng.uniform(np.log(64), np.log(8192)))))
        n = int(rng.integers(max(1, m // 8), max(2, m // 2) + 1))
        q = int(rng.integers(1, 17))
        cond = float(np.exp(rng.uniform(np.log(1e2), np.log(1e8))))
        a = conditioned_dev(R, torch, h, m, n, cond, 100 + 3 * i)
        xt = torch.from_numpy(rng.standard_normal((n, q))).cuda()
        b = a @ xt + 1e-8 * torch.from_numpy(rng.standard_normal((m, q))).cuda()
        reqs.append((a.cpu().numpy(), b.cpu().numpy(), cond))
    gaps = rng.exponential(1.0 / rate, count)
    srv = R.Server(hbm_budget=64 << 30, workers=4, mode=mode, max_iter=30, tol=1e-13)
    t0 = time.monotonic() + 0.05
    arrivals, ids = [], []
    t_next = t0
    for (a, b, _), g in zip(reqs, gaps):
        t_next += g
        d = t_next - time.monotonic()
        if d > 0:
            time.sleep(d)
        arrivals.append(t_next)
        ids.append(srv.submit(a, b)[1])
    lat, worst, paths = [], 0.0, {}
    lag, wait, service, by_path, flops = [], [], [], {}, []
    t_end = 0.0
    for (a, b, _), arr, rid in zip(reqs, arrivals, ids):
        x, info = srv.poll(rid)
        lat.append(info.t_done - arr)
        lag.append(info.t_submit - arr)
        wait.append(info.t_admit - info.t_submit)
        service.append(info.t_done - info.t_admit)
        by_path.setdefault(R.PATHS[info.path], []).append(info.t_done - arr)
        mm, nn = a.shape
        flops.append(2.0 * mm * nn * nn - 2.0 * nn ** 3 / 3.0)
        t_end = max(t_end, info.t_done)
        paths[R.PATHS[info.path]] = paths.get(R.PATHS[info.path], 0) + 1
        r = b - a @ x
        an = np.linalg.norm(a)
        e = np.max(np.linalg.norm(a.T @ r, axis=0) / (an * (an * np.linalg.norm(x, axis=0) + np.linalg.norm(b, axis=0))))
        worst = max(worst, float(e))
    stats = srv.stats()
    srv.close()
    lat = np.array(lat) * 1e3
    pct = lambda v, q: round(float(np.percentile(np.asarray(v) * 1e3, q)), 3)
    # where the tail comes from: admission wait vs time on the device, per
    # path, and the largest tenth of the requests (by factorization flops)
    big = np.asarray(flops) >= np.percentile(flops, 90)
    breakdown = {
        # scheduled arrival -> submit called (submit copies A and b into
        # the request, so a large request delays the client loop's next
        # submissions); the copy itself counts in the admission wait
        "submit_lag_ms": {"p50": pct(lag, 50), "p99": pct(lag, 99)},
        "admission_wait_ms": {"p50": pct(wait, 50), "p99": pct(wait, 99)},
        "service_ms": {"p50": pct(service, 50), "p99": pct(service, 99)},
        "by_path_ms": {k: {"count": len(v), "p50": pct(v, 50), "p99": pct(v, 99)} for k, v in by_path.items()},
        "largest_tenth_ms": {"count": int(big.sum()), "p50": round(float(np.percentile(lat[big], 50)), 3),
                             "max": round(float(lat[big].max()), 3)},
        "rest_ms": {"p50": round(float(np.percentile(lat[~big], 50)), 3),
                    "p99": round(float(np.percentile(lat[~big], 99)), 3)},
    }
    return {"requests": count, "rate_per_s": rate, "mode": "mixed+fp64-fallback" if mode == 1 else "fp64",
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "throughput_rps": round(count / (t_end - arrivals[0]), 2),
            "max_berr": worst, "paths": paths, "rejected": stats["rejected"], "batches": stats["batches"],
            "breakdown": breakdown}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nb", type=int, default=0)
    ap.add_argument("--rpt", type=int, default=0)
    ap.add_argument("--no-lookahead", action="store_true", help="serialize panels and updates")
    ap.add_argument("--panel-cus", type=int, default=0,
                    help="CUs reserved for the look-ahead panel stream (0 automatic, -1 none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--legs", default="all",
                    help="comma list of extra legs: complex,tall,batched,mixed,serving,lq; 'none' or 'all'")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding of this path is a design note, not a measured run.", file=sys.stderr)
        sys.exit(2)

    import torch
    import rnp2_amd as R

    m, n = args.m, args.n
    seed = [1, 2, 3, 5]
    h = R.Handle()
    if args.nb or args.rpt or args.no_lookahead or args.panel_cus:
        h.set_tuning(nb=args.nb, panel_rpt=args.rpt, lookahead=-1 if args.no_lookahead else 0,
                     panel_cus=args.panel_cus)
    a0 = R.colmajor_empty(m, n, torch.float64)
    h.larnv(2, seed, a0)          # RNP Normal01 stream, column-major fill
    a = R.colmajor_empty(m, n, torch.float64)
    tau = torch.empty(min(m, n), dtype=torch.float64, device="cuda")
    h.synchronize()
    stream = torch.cuda.ExternalStream(h.stream)

    def step():
        with torch.cuda.stream(stream):
            a.copy_(a0)
        h.geqrf(a, tau)

    for _ in range(args.warmup):
        step()
    h.synchronize()
    torch.cuda.synchronize()

    # Timed steps: only the roofline kernel (the trailing update) carries
    # HIP events, so that the instrumentation adds as little as possible.
    h.profile(True, kernels=["update"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    h.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    upd = h.profile_read("update")
    # Per-family breakdown from separate, untimed steps with every family
    # instrumented.
    bsteps = min(args.steps, 2)
    h.profile(True)
    for _ in range(bsteps):
        step()
    h.synchronize()
    wg = h.profile_read("wgemm")
    pan = h.profile_read("panel")
    tf = h.profile_read("tfactor")
    upd_b = h.profile_read("update")
    h.profile(False)

    # Correctness of the timed factorization (outside the timed region), on
    # the device: Q from GenerateQ, ||A - QR|| / (n eps ||A||) and
    # ||Q^T Q - I|| / (n eps), both required <= 10.
    eps = float(np.finfo(np.float64).eps)
    k = min(m, n)
    q = R.colmajor_empty(m, k, torch.float64)
    with torch.cuda.stream(stream):
        q.copy_(a[:, :k])
    h.orgqr(q, tau, n=k)
    h.synchronize()
    r = torch.triu(a[:k, :])
    res = float(torch.linalg.norm(a0 - q @ r) / (k * eps * torch.linalg.norm(a0)))
    orth = float(torch.linalg.norm(q.t() @ q - torch.eye(k, dtype=torch.float64, device="cuda")) / (k * eps))
    del q, r
    ok = bool(res <= 10 and orth <= 10)

    fl = qr_flops(m, n)
    value = fl / (ms * 1e-3) / 1e9
    achieved = (upd["flops"] / (upd["ms"] * 1e-3) / 1e12) if upd["ms"] > 0 else 0.0
    line = {
        "metric": "FP64 QR GFLOP/s at 8192^2 (QR::Factor)",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FP64 QR::Factor of a {m}x{n} column-major matrix, N(0,1) entries "
                               f"(RNP Normal01 stream, iseed {seed}), resident in HBM",
                   "m": m, "n": n, "nb": args.nb or 64, "flops_per_step": fl},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4),
                     "traffic": update_traffic_bytes() if (m, n) == (8192, 8192) else None,
                     "algorithmic_bytes_per_launch": round(upd["bytes"] / max(1, upd["launches"])),
                     "traffic_source": "profiles/qr_factor_8192_kernels.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE)",
                     "kernel": "rnpx::dg::dgemm_kernel CINIT (trailing update C -= V Z, FP64 MFMA, "
                               "rnp2_amd/csrc/rnpx_dgemm.hip)",
                     "launches": upd["launches"], "kernel_ms_per_step": round(upd["ms"] / args.steps, 3)},
        "breakdown_ms_per_step": {
            "steps": bsteps,
            "update_gemm": round(upd_b["ms"] / bsteps, 3),
            "w_gemm": round(wg["ms"] / bsteps, 3),
            "panel": round(pan["ms"] / bsteps, 3),
            "tfactor": round(tf["ms"] / bsteps, 3),
        },
        "panel_hbm_gbs": round(pan["bytes"] / (pan["ms"] * 1e-3) / 1e9, 1) if pan["ms"] > 0 else None,
        "check": {"ok": ok, "residual": round(res, 4), "orthogonality": round(orth, 4)},
    }
    # Each extra leg runs on a handle of its own, as in a process of its own:
    # the headline's handle (two streams) is closed first, so that the legs'
    # streams do not share the runtime's few hardware queues with idle ones.
    h.synchronize()
    h.close()
    del a, a0, tau
    torch.cuda.synchronize()
    legs = [] if args.legs == "none" else (["complex", "tall", "batched", "mixed", "serving", "lq", "qlrq", "lu"]
                                           if args.legs == "all" else args.legs.split(","))
    extra = {}
    for leg in legs:
        print(f"[bench] leg {leg}", file=sys.stderr, flush=True)
        try:
            if leg == "complex":
                extra["complex_4096"] = leg_factor(R, torch, 4096, 4096, True)
            elif leg == "tall":
                extra["tall_32768x4096"] = leg_factor(R, torch, 32768, 4096, False)
            elif leg == "batched":
                extra["batched_10k_ls"] = {"d": leg_batched(R, torch, False), "z": leg_batched(R, torch, True)}
            elif leg == "mixed":
                extra["mixed_ls_16384x2048"] = leg_mixed_ls(R, torch)
            elif leg == "serving":
                extra["serving"] = leg_serving(R, torch)
            elif leg == "lq":
                extra["lq_4096x8192"] = leg_lq(R, torch)
            elif leg == "qlrq":
                extra["ql_rq_8192x4096"] = leg_qlrq(R, torch)
            elif leg == "lu":
                extra["lu_8192"] = leg_lu(R, torch)
        except Exception as e:  # a failing leg is reported, the headline line still prints
            extra[leg] = {"error": repr(e)}
    if extra:
        line["extra"] = extra
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(m, n, seed, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
