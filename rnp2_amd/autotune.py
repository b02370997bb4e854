"""Autotuner for the device QR: times candidate configurations on the GPU and
writes the table librnpx.so reads (rnpx_tuning.json beside the library, or
the path given), the device counterpart of RNP's QR::Tuning hooks
(RNP/LA/QR.hpp:41-51), which fix nb = 32 for every shape.

Factor: for each shape and element type, every outer block width in
{32, 64, 128, 256} (the width of the compact-WY block the trailing matrix sees),
W = V^H C split factor and CU count reserved for the look-ahead panel stream
in the candidates, on RNP Normal01 input resident
in HBM, best of `reps` timed runs after a warm-up; the fastest is recorded.
Batched: for the batched configuration's distribution (m log-uniform in
[16, 256], n uniform in [8, m]), every small-tier LDS budget per one-wave
team and large-tier blocked-path column threshold in the candidates.

    python -m rnp2_amd.autotune                       # the benchmarked shapes
    python -m rnp2_amd.autotune --shapes 4096x4096 --out /tmp/t.json
"""
import argparse
import json
import os
import time

import numpy as np

from . import lib, Handle, colmajor_empty, make_batch, device_name

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "rnpx_tuning.json")
WIDTHS = (32, 64, 128, 256)
KSPLITS = (0, 4, 8, 16)
PANEL_CUS = (0, 32, 64)  # CUs reserved for the look-ahead panel stream (0: shared)
BATCH_LDS = (20480, 40960, 81920)
BATCH_BLOCKED_MIN = (17, 33, 65)


def _time(fn, sync, reps):
    fn()
    sync()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        sync()
        best = min(best, time.perf_counter() - t0)
    return best


def tune_factor(m, n, cplx=False, widths=WIDTHS, ksplits=KSPLITS, panel_cus=PANEL_CUS, reps=3):
    """Times geqrf of an m x n matrix for every (width, split, panel CUs);
    returns the fastest entry and all timings."""
    import torch
    dt = torch.complex128 if cplx else torch.float64
    h = Handle()
    a0 = colmajor_empty(m, n, dt)
    h.larnv(2, [1, 2, 3, 5], a0)
    a = colmajor_empty(m, n, dt)
    tau = torch.empty(min(m, n), dtype=dt, device="cuda")
    h.synchronize()
    stream = torch.cuda.ExternalStream(h.stream)
    results = []
    for w in widths:
        for ks in ksplits:
            for pc in panel_cus:
                h.set_tuning(nb=w, ksplit=ks, panel_cus=pc if pc else -1)

                def run():
                    with torch.cuda.stream(stream):
                        a.copy_(a0)
                    h.geqrf(a, tau)

                t = _time(run, h.synchronize, reps)
                results.append({"nb": w, "ksplit": ks, "panel_cus": pc, "ms": round(t * 1e3, 3)})
    h.set_tuning()
    best = pick_factor(results)
    entry = {"kind": "factor", "dtype": "z" if cplx else "d", "m": m, "n": n, "nb": best["nb"],
             "ksplit": best["ksplit"], "panel_cus": best["panel_cus"], "ms": best["ms"]}
    return entry, results


def pick_factor(results, margin=0.01):
    """The fastest candidate, except that a shared-CU one (panel_cus 0)
    within `margin` of it is preferred: a CU partition has to pay for itself."""
    best = min(results, key=lambda r: r["ms"])
    shared = [r for r in results if r.get("panel_cus", 0) == 0]
    if shared:
        bs = min(shared, key=lambda r: r["ms"])
        if bs["ms"] <= best["ms"] * (1.0 + margin):
            return bs
    return best


def tune_batched(count=10000, cplx=False, lds=BATCH_LDS, blocked_min=BATCH_BLOCKED_MIN, reps=3, seed=0):
    """Times the batched factorization of `count` problems of the batched
    configuration for every (LDS per team, blocked threshold)."""
    import torch
    rng = np.random.default_rng(seed)
    ms = np.rint(np.exp(rng.uniform(np.log(16), np.log(256), count))).astype(int)
    ns = np.array([rng.integers(8, mm + 1) for mm in ms])
    shapes = list(zip(ms.tolist(), ns.tolist()))
    desc, na, nt, nbb = make_batch(shapes, nrhs=1)
    dt = torch.complex128 if cplx else torch.float64
    A0 = torch.randn(na, dtype=torch.float64, device="cuda")
    if cplx:
        A0 = torch.complex(A0, torch.randn(na, dtype=torch.float64, device="cuda"))
    A = torch.empty_like(A0)
    B0 = torch.ones(max(nbb, 1), dtype=dt, device="cuda")
    B = torch.empty_like(B0)
    tau = torch.empty(max(nt, 1), dtype=dt, device="cuda")
    results = []
    path = None
    import tempfile
    for L in lds:
        for bm in blocked_min:
            # a one-entry table selects the candidate for this call
            fd, path = tempfile.mkstemp(suffix=".json")
            with os.fdopen(fd, "w") as f:
                json.dump({"version": 1, "entries": [{"kind": "batched", "dtype": "z" if cplx else "d", "lds": L,
                                                      "blocked_min": bm}]}, f)
            lib().rnpx_tuning_load(path.encode())
            h = Handle()

            def run():
                A.copy_(A0)
                B.copy_(B0)
                torch.cuda.synchronize()
                h.geqrf_batched(desc, A, tau, B)

            t = _time(run, h.synchronize, reps)
            results.append({"lds": L, "blocked_min": bm, "ms": round(t * 1e3, 3)})
            h.close()
            os.unlink(path)
    lib().rnpx_tuning_load(None)
    best = min(results, key=lambda r: r["ms"])
    entry = {"kind": "batched", "dtype": "z" if cplx else "d", "lds": best["lds"], "blocked_min": best["blocked_min"],
             "ms": best["ms"]}
    return entry, results


def write_table(entries, path=DEFAULT_PATH, extra=None):
    doc = {"version": 1, "device": device_name(), "entries": entries}
    if extra:
        doc.update(extra)
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--shapes", default="8192x8192,32768x4096,16384x2048", help="real factor shapes, MxN comma list")
    ap.add_argument("--cshapes", default="4096x4096", help="complex factor shapes")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args(argv)
    entries, log = [], {}
    for spec, cplx in [(args.shapes, False), (args.cshapes, True)]:
        for s in [x for x in spec.split(",") if x]:
            m, n = (int(v) for v in s.lower().split("x"))
            e, res = tune_factor(m, n, cplx, reps=args.reps)
            entries.append(e)
            log[f"factor {'z' if cplx else 'd'} {m}x{n}"] = res
            print(json.dumps(e), flush=True)
    if not args.no_batched:
        for cplx in (False, True):
            e, res = tune_batched(cplx=cplx, reps=args.reps)
            entries.append(e)
            log[f"batched {'z' if cplx else 'd'}"] = res
            print(json.dumps(e), flush=True)
    # the table is written for a fresh process; the timings go beside it
    write_table(entries, args.out, {"candidates": log})
    lib().rnpx_tuning_load(args.out.encode())
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
