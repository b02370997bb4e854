"""GPU parity of QL (RNP/LA/QL.hpp) and RQ (RNP/LA/RQ.hpp): Factor, MultQ,
GenerateQ against the oracle's restatements on the same seeded inputs.

Bar (the QR one): the triangular factor and tau elementwise (column-scaled
for QL's L, row-scaled for RQ's R) <= 1e-11, the stored reflector entries
<= 1e-9 absolute, ||A - Q L|| (||A - R Q||) / (n eps ||A||) <= 10 and the
orthogonality error / (n eps) <= 10; MultQ and GenerateQ <= 1e-11 relative.
"""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu

TOL = 1e-11


def seeded(m, n, cplx, seed):
    iseed = np.array([seed % 4096, m % 4096, n % 4096, 7], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed, complex_=cplx)


def ql_mask(m, n):
    """True on L: the lower triangle ending at (m-1, n-1) and, when n > m,
    the leading n-m columns."""
    k = min(m, n)
    i = np.arange(m)[:, None]
    j = np.arange(n)[None, :]
    return (j < n - k) | ((i >= m - k) & ((j - (n - k)) <= (i - (m - k))))


def rq_mask(m, n):
    """True on R: the upper triangle ending at (m-1, n-1) and, when m > n,
    the leading m-n rows."""
    k = min(m, n)
    i = np.arange(m)[:, None]
    j = np.arange(n)[None, :]
    return (i < m - k) | ((j - (n - k)) >= (i - (m - k)))


def check_ql(a, tag):
    print(f"[case] {tag} shape={a.shape} dtype={a.dtype}", flush=True)
    m, n = a.shape
    af_o, tau_o = O.ql_factor(a)
    af_g, tau_g = R.ql_factor(a)
    mk = ql_mask(m, n)
    e_l = rel_err_colscaled(np.where(mk, af_g, 0), np.where(mk, af_o, 0))
    e_t = float(np.max(np.abs(tau_g - tau_o))) if tau_o.size else 0.0
    e_v = float(np.max(np.abs(np.where(mk, 0, af_g - af_o)))) if af_o.size else 0.0
    out = {"L": e_l, "tau": e_t, "V": e_v}
    if m >= n and n:
        Q = R.ql_generateq(af_g, tau_g)
        L = np.tril(af_g[m - n:, :])
        out["res"] = qr_residual(a, Q, L)
        out["orth"] = orth_error(Q)
    print(f"        {out}", flush=True)
    assert e_l <= TOL and e_t <= TOL and e_v <= 1e-9, out
    if "res" in out:
        assert out["res"] <= 10 and out["orth"] <= 10, out
    return af_g, tau_g


def check_rq(a, tag):
    print(f"[case] {tag} shape={a.shape} dtype={a.dtype}", flush=True)
    m, n = a.shape
    af_o, tau_o = O.rq_factor(a)
    af_g, tau_g = R.rq_factor(a)
    mk = rq_mask(m, n)
    e_r = rel_err_colscaled(np.where(mk, af_g, 0).T, np.where(mk, af_o, 0).T)  # row-scaled
    e_t = float(np.max(np.abs(tau_g - tau_o))) if tau_o.size else 0.0
    e_v = float(np.max(np.abs(np.where(mk, 0, af_g - af_o)))) if af_o.size else 0.0
    out = {"R": e_r, "tau": e_t, "V": e_v}
    if m <= n and m:
        Q = R.rq_generateq(af_g, tau_g)
        Rm = np.triu(af_g[:, n - m:])
        out["res"] = qr_residual(a.conj().T, Q.conj().T, Rm.conj().T)
        out["orth"] = orth_error(Q.conj().T)
    print(f"        {out}", flush=True)
    assert e_r <= TOL and e_t <= TOL and e_v <= 1e-9, out
    if "res" in out:
        assert out["res"] <= 10 and out["orth"] <= 10, out
    return af_g, tau_g


SHAPES = [(160, 170), (512, 256), (256, 512), (1, 1), (1, 7), (7, 1), (300, 1000), (1000, 300), (2048, 2048)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_ql_factor_matches_oracle(shape, cplx):
    m, n = shape
    check_ql(seeded(m, n, cplx, 41), f"ql cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_rq_factor_matches_oracle(shape, cplx):
    m, n = shape
    check_rq(seeded(m, n, cplx, 43), f"rq cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
def test_ql_rq_edges(cplx):
    """Re(alpha) == 0 (beta > 0), a zero column / row (tau = 0), tiny and huge
    columns / rows."""
    a = seeded(90, 40, cplx, 9)
    a[89, 39] = 0.0
    a[:, 30] = 0.0
    a[:, 20] *= 1e-300
    a[:, 10] *= 1e300
    af, tau = check_ql(a, f"ql edges cplx={cplx}")
    assert af[89, 39].real > 0 and tau[30] == 0
    b = seeded(40, 90, cplx, 10)
    b[39, 89] = 0.0
    b[30, :] = 0.0
    b[20, :] *= 1e-300
    b[10, :] *= 1e300
    bf, tb = check_rq(b, f"rq edges cplx={cplx}")
    assert bf[39, 89].real > 0 and tb[30] == 0


@pytest.mark.parametrize("fam", ["ql", "rq"])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_multq(fam, side, trans, cplx):
    print(f"[case] {fam} multq {side}{trans} cplx={cplx}", flush=True)
    rng = np.random.default_rng(5)
    for k, nq in ((20, 40), (150, 400)):
        if fam == "ql":
            af, tau = O.ql_factor(seeded(nq, k, cplx, k))
        else:
            af, tau = O.rq_factor(seeded(k, nq, cplx, k))
        c = rng.standard_normal((nq, 33) if side == "L" else (33, nq))
        if cplx:
            c = c + 1j * rng.standard_normal(c.shape)
        want = (O.ql_multq if fam == "ql" else O.rq_multq)(side, trans, af, tau, c, k=k)
        got = (R.ql_multq if fam == "ql" else R.rq_multq)(side, trans, af, tau, c, k=k)
        err = float(np.max(np.abs(got - want)) / max(np.abs(want).max(), 1e-300))
        print(f"        k={k} nq={nq}: {err:.2e}", flush=True)
        assert err <= TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,k", [(100, 60, 60), (100, 60, 40), (300, 300, 300)])
def test_generateq(m, n, k, cplx):
    """QL: a (m x n) with the reflectors in its last k columns; RQ: the
    transposed shapes with the reflectors in the last k rows."""
    print(f"[case] ql/rq generateq {m}x{n} k={k} cplx={cplx}", flush=True)
    af, tau = O.ql_factor(seeded(m, k, cplx, 11))
    a = np.zeros((m, n), dtype=af.dtype, order="F")
    a[:, n - k:] = af
    want = O.ql_generateq(a.copy(), tau, k=k)
    got = R.ql_generateq(a.copy(), tau, k=k)
    err = float(np.max(np.abs(got - want)))
    print(f"        ql {err:.2e}", flush=True)
    assert err <= TOL
    bf, tb = O.rq_factor(seeded(k, m, cplx, 12))
    b = np.zeros((n, m), dtype=bf.dtype, order="F")
    b[n - k:, :] = bf
    want = O.rq_generateq(b.copy(), tb, k=k)
    got = R.rq_generateq(b.copy(), tb, k=k)
    err = float(np.max(np.abs(got - want)))
    print(f"        rq {err:.2e}", flush=True)
    assert err <= TOL


def test_device_handle_matches_dropin():
    """Handle (device pointers) and drop-in (host pointers) give the same
    factors bit for bit."""
    import torch
    print("[case] ql/rq device handle", flush=True)
    h = R.Handle()
    for fam, shape in (("ql", (700, 300)), ("rq", (300, 700))):
        a = seeded(*shape, False, 2)
        af_h, tau_h = (R.ql_factor if fam == "ql" else R.rq_factor)(a)
        d = R.colmajor_empty(*shape, torch.float64)
        d.copy_(torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda").t())
        tau = torch.zeros(min(shape), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        (h.geqlf if fam == "ql" else h.gerqf)(d, tau)
        h.synchronize()
        assert np.array_equal(d.cpu().numpy(), af_h) and np.array_equal(tau.cpu().numpy(), tau_h), fam
