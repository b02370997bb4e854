"""GPU parity of LQ (RNP/LA/LQ.hpp: Factor, MultQ, GenerateQ) against the
oracle's row-wise restatement, on the same seeded inputs.

Bar (the QR one, transposed): L and tau elementwise (L row-scaled) <= 1e-11,
the conjugated reflector rows <= 1e-9 absolute, ||A - L Q||/(n eps ||A||) <= 10
and ||Q Q^H - I||/(n eps) <= 10; MultQ and GenerateQ <= 1e-11 relative.
"""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu

TOL = 1e-11


def seeded(m, n, cplx, seed):
    iseed = np.array([seed % 4096, m % 4096, n % 4096, 5], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed, complex_=cplx)


def check_lq(a, tag):
    print(f"[case] {tag} shape={a.shape} dtype={a.dtype}", flush=True)
    m, n = a.shape
    k = min(m, n)
    af_o, tau_o = O.lq_factor(a)
    af_g, tau_g = R.lq_factor(a)
    L_o, L_g = np.tril(af_o)[:, :k], np.tril(af_g)[:, :k]
    e_l = rel_err_colscaled(L_g.conj().T, L_o.conj().T)      # row-scaled on L
    e_t = float(np.max(np.abs(tau_g - tau_o))) if k else 0.0
    e_v = float(np.max(np.abs(np.triu(af_g, 1) - np.triu(af_o, 1)))) if k else 0.0
    out = {"L": e_l, "tau": e_t, "V": e_v}
    if k:
        Q = R.lq_generateq(af_g[:k], tau_g, m=k)
        out["res"] = qr_residual(a.conj().T, Q.conj().T, L_g.conj().T)
        out["orth"] = orth_error(Q.conj().T)
    print(f"        {out}", flush=True)
    assert e_l <= TOL and e_t <= TOL and e_v <= 1e-9, out
    if k:
        assert out["res"] <= 10 and out["orth"] <= 10, out
    return af_g, tau_g


SHAPES = [(160, 170), (512, 256), (256, 512), (1, 1), (1, 7), (7, 1), (300, 1000), (1000, 300), (2048, 2048)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_lq_factor_matches_oracle(shape, cplx):
    m, n = shape
    check_lq(seeded(m, n, cplx, 31), f"lq cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
def test_lq_edge_rows(cplx):
    """Re(alpha) == 0 (beta > 0), a zero row (tau = 0), tiny and huge rows."""
    a = seeded(40, 90, cplx, 7)
    a[0, 0] = 0.0
    a[3, :] = 0.0
    a[5, :] *= 1e-300
    a[6, :] *= 1e300
    af, tau = check_lq(a, f"lq edges cplx={cplx}")
    assert af[0, 0].real > 0 and tau[3] == 0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_lq_multq(side, trans, cplx):
    print(f"[case] lq multq {side}{trans} cplx={cplx}", flush=True)
    rng = np.random.default_rng(3)
    for k, nq in ((20, 40), (150, 400)):
        a = seeded(k, nq, cplx, k)
        af, tau = O.lq_factor(a)
        c = rng.standard_normal((nq, 33) if side == "L" else (33, nq))
        if cplx:
            c = c + 1j * rng.standard_normal(c.shape)
        want = O.lq_multq(side, trans, af, tau, c, k=k)
        got = R.lq_multq(side, trans, af, tau, c, k=k)
        err = float(np.max(np.abs(got - want)) / max(np.abs(want).max(), 1e-300))
        print(f"        k={k} nq={nq}: {err:.2e}", flush=True)
        assert err <= TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,k", [(60, 100, 60), (60, 100, 40), (300, 300, 300)])
def test_lq_generateq(m, n, k, cplx):
    print(f"[case] lq generateq {m}x{n} k={k} cplx={cplx}", flush=True)
    a = seeded(m, n, cplx, 11)
    af, tau = O.lq_factor(a)
    want = O.lq_generateq(af, tau[:k], m=m, k=k)
    got = R.lq_generateq(af, tau[:k], m=m, k=k)
    err = float(np.max(np.abs(got - want)))
    print(f"        {err:.2e}", flush=True)
    assert err <= TOL


def test_lq_device_handle_matches_dropin():
    """Handle (device pointers) and drop-in (host pointers) give the same
    factors bit for bit."""
    import torch
    print("[case] lq device handle", flush=True)
    a = seeded(300, 700, False, 2)
    af_h, tau_h = R.lq_factor(a)
    h = R.Handle()
    d = R.colmajor_empty(300, 700, torch.float64)
    d.copy_(torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda").t())
    tau = torch.zeros(300, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    h.gelqf(d, tau)
    h.synchronize()
    assert np.array_equal(d.cpu().numpy(), af_h) and np.array_equal(tau.cpu().numpy(), tau_h)
