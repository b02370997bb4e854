"""The allocation-free batched entry point (rnpx_dgeqrf_batched_noalloc)
against the regular one (rnpx_dgeqrf_batched), through the C ABI.

Both run the same kernels on the same problems, so the factors, tau and the
least-squares solutions must agree bit for bit.  Two calls in a row, the
second with fewer problems (the shared static scratch keeps the first call's
contents).  The handle's own workspace must not grow across the calls, and a
count above the static maximum is refused before anything is touched.
"""
import numpy as np
import pytest

import rnp2_amd as R

pytestmark = pytest.mark.gpu


def c3_shapes(count, seed):
    rng = np.random.default_rng(seed)
    m = np.rint(np.exp(rng.uniform(np.log(16), np.log(256), count))).astype(int)
    n = np.array([rng.integers(8, mm + 1) for mm in m])
    return list(zip(m.tolist(), n.tolist()))


def _run(shapes, seed):
    import torch
    desc, na, nt, nb_ = R.make_batch(shapes, nrhs=2)
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(na)
    b = rng.standard_normal(max(nb_, 1))
    out = []
    for fn in ("regular", "noalloc"):
        A = torch.from_numpy(a).cuda()
        B = torch.from_numpy(b).cuda()
        tau = torch.zeros(max(nt, 1), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        if fn == "regular":
            h = R.Handle()
            h.geqrf_batched(desc, A, tau, B)
            h.synchronize()
        else:
            h = _NOALLOC_HANDLE
            before = h.workspace_bytes()
            st = h.geqrf_batched_noalloc(desc, A, tau, B)
            assert st == 0, st
            h.synchronize()
            assert h.workspace_bytes() == before
        out.append((A.cpu().numpy(), tau.cpu().numpy(), B.cpu().numpy()))
    return out


_NOALLOC_HANDLE = None


def test_noalloc_matches_regular_two_calls():
    global _NOALLOC_HANDLE
    _NOALLOC_HANDLE = R.Handle()
    big = c3_shapes(600, 11) + [(256, 256), (256, 200), (200, 180), (17, 8), (1, 1), (5, 9)]
    small = c3_shapes(40, 12) + [(250, 240)]
    for tag, shapes, seed in (("first", big, 1), ("second, smaller", small, 2)):
        print(f"[case] noalloc batched {tag}: {len(shapes)} problems", flush=True)
        (a0, t0, b0), (a1, t1, b1) = _run(shapes, seed)
        assert np.array_equal(a0, a1) and np.array_equal(t0, t1) and np.array_equal(b0, b1), tag
        assert np.all(np.isfinite(t1))


def test_noalloc_refuses_oversize():
    import torch
    h = R.Handle()
    shapes = [(16, 8)] * 16385
    desc, na, nt, _ = R.make_batch(shapes)
    A = torch.ones(na, dtype=torch.float64, device="cuda")
    tau = torch.zeros(nt, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = h.geqrf_batched_noalloc(desc, A, tau)
    h.synchronize()
    assert st == R.RNPX_ERR_UNSUPPORTED
    assert bool((A == 1).all()) and bool((tau == 0).all())
