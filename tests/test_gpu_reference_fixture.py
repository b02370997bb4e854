"""GPU results against the real reference's own outputs.

tests/golden/rnp_reference_qr.npz holds RNP's blocked QR (RNP/LA/QR.hpp
Factor, MultQ, GenerateQ) built with its src/BLAS_mix.cpp against MKL
(written by tests/golden/make_golden_qr.py).  The GPU path goes through the
C ABI (the drop-in rnpx_{d,z}geqrf / ormqr / orgqr, and the batched entry
point for the batched configuration's problems).  Tolerance: column-scaled
elementwise error <= 1e-11 for R, V and tau (the GPU blocks and sums in a
different order from RNP + MKL); the sign rule and tau = 0 decisions exact.
"""
import os

import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "rnp_reference_qr.npz"))
TOL = 1e-11


def rnp_stream(m, n, cplx):
    O.reset_internal_seed()
    return O.generate_matrix(O.UNIFORM_11, m, n, None, cplx)


def seeded(m, n, p):
    iseed = np.array([p % 4096, (7 * p + 1) % 4096, (13 * p + 2) % 4096, (2 * p + 1) % 4096], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed)


def trailing_rank(af):
    k = min(af.shape)
    d = np.abs(np.diag(af[:k, :k]))
    tiny = d <= 1e-10 * max(np.max(d), 1e-300) if k else d
    r = k
    while r > 0 and tiny[r - 1]:
        r -= 1
    return r


def errors(af, tau, af_r, tau_r):
    """Column-scaled errors of R, V and tau up to the numerical rank (past it
    the reflectors are those of rounding noise)."""
    r = trailing_rank(af_r)
    e_r = rel_err_colscaled(np.triu(af[:r]), np.triu(af_r[:r])) if r else 0.0
    vr, vg = np.tril(af_r, -1)[:, :r], np.tril(af, -1)[:, :r]
    vs = np.max(np.abs(vr), axis=0) if vr.size else np.ones(r)
    vs[vs == 0] = 1.0
    e_v = float(np.max(np.abs(vg - vr) / vs)) if vr.size else 0.0
    e_t = float(np.max(np.abs(tau[:r] - tau_r[:r]))) if r else 0.0
    return e_r, e_v, e_t


def test_config1_against_reference():
    print("[case] reference fixture config 1 512x256", flush=True)
    a = rnp_stream(512, 256, False)
    af, tau = R.factor(a)
    e = errors(af, tau, G["c1_afac"], G["c1_tau"])
    q = R.generateq(af, tau, n=256)
    e_q = float(np.max(np.abs(q - G["c1_q"])))
    res, orth = qr_residual(a, q, np.triu(af[:256])), orth_error(q)
    print(f"        R/V/tau {e} Q {e_q:.2e} res {res:.3f} orth {orth:.3f}", flush=True)
    assert max(e) <= TOL and e_q <= TOL and res <= 10 and orth <= 10


@pytest.mark.parametrize("key,m,n,cplx", [("reft_d_170x160", 170, 160, False), ("reft_d_166x170", 166, 170, False),
                                          ("reft_z_170x160", 170, 160, True)])
def test_reference_test_shapes(key, m, n, cplx):
    print(f"[case] reference fixture {key}", flush=True)
    af, tau = R.factor(rnp_stream(m, n, cplx))
    e = errors(af, tau, G[key + "_afac"], G[key + "_tau"])
    print(f"        {e}", flush=True)
    assert max(e) <= TOL


@pytest.mark.parametrize("pre", ["d", "z"])
def test_edge_matrices(pre):
    for i in range(int(G[f"edge_{pre}_count"])):
        a = G[f"edge_{pre}_{i}_a"]
        print(f"[case] reference fixture edge {pre} {i} {a.shape}", flush=True)
        af, tau = R.factor(a)
        af_r, tau_r = G[f"edge_{pre}_{i}_afac"], G[f"edge_{pre}_{i}_tau"]
        e = errors(af, tau, af_r, tau_r)
        print(f"        {e}", flush=True)
        assert max(e) <= TOL, (i, e)
        r = trailing_rank(af_r)
        assert np.array_equal(tau[:r] == 0, tau_r[:r] == 0)
        d, d_r = np.diag(af[:r, :r]), np.diag(af_r[:r, :r])
        assert np.array_equal(np.sign(d.real), np.sign(d_r.real))


@pytest.mark.parametrize("pre", ["d", "z"])
def test_multq(pre):
    print(f"[case] reference fixture MultQ {pre}", flush=True)
    for side in "LR":
        for trans in "NC":
            c = G[f"mq_{pre}_{side}{trans}_c"]
            out = R.multq(side, trans, G[f"mq_{pre}_afac"], G[f"mq_{pre}_tau"], c)
            ref = G[f"mq_{pre}_{side}{trans}_out"]
            e = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
            print(f"        {side}{trans} {e:.2e}", flush=True)
            assert e <= TOL


def test_short_workspace():
    """A short workspace changes the reference's blocking; the drop-in's
    result is the same factorization to rounding."""
    print("[case] reference fixture short workspace 200x140", flush=True)
    a = seeded(200, 140, 999)
    af, tau = R.factor(a)
    e = errors(af, tau, G["short_afac"], G["short_tau"])
    print(f"        {e}", flush=True)
    assert max(e) <= TOL


def test_batched_problems():
    """Problems drawn from the batched configuration, in one batched launch."""
    import torch
    shapes = [tuple(s) for s in G["batch_shapes"].tolist()]
    print(f"[case] reference fixture batched {shapes}", flush=True)
    desc, na, nt, _ = R.make_batch(shapes)
    a = np.zeros(na)
    for i, (m, n) in enumerate(shapes):
        o = int(desc[i]["a_off"])
        a[o:o + m * n] = seeded(m, n, i).T.reshape(-1)
    A = torch.from_numpy(a).cuda()
    tau = torch.zeros(max(nt, 1), dtype=torch.float64, device="cuda")
    h = R.Handle()
    h.geqrf_batched(desc, A, tau)
    h.synchronize()
    Ag, tg = A.cpu().numpy(), tau.cpu().numpy()
    for i, (m, n) in enumerate(shapes):
        o, t = int(desc[i]["a_off"]), int(desc[i]["tau_off"])
        af = Ag[o:o + m * n].reshape(n, m).T
        e = errors(af, tg[t:t + min(m, n)], G[f"batch_{i}_afac"], G[f"batch_{i}_tau"])
        assert max(e) <= TOL, (i, m, n, e)
