"""Complex 32-column panels by CholeskyQR2 and Householder reconstruction
(rnpx_zcqr.hip) against the oracle's QR::Factor (QR.hpp:129-197): accepted
panels, and panels the path must reject (a column 1e-20 times the others, a
zero column, a repeated column, a column of 1e200) -- those are factored by
the column-wise kernel the same launch sequence gates on the fail flag.
Tolerances: R and tau column-scaled elementwise error <= 1e-11 against the
oracle, ||A - QR|| / (n eps ||A||) <= 10, ||Q^H Q - I|| / (n eps) <= 10."""
import numpy as np
import pytest
import torch

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu


def cmat(m, n, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))


def factor_dev(a, panel=0):
    m, n = a.shape
    h = R.Handle()
    try:
        h.set_tuning(nb=32, panel=panel)
        A = R.colmajor_empty(m, n, torch.complex128)
        A.copy_(torch.from_numpy(np.ascontiguousarray(a)).to("cuda"))
        tau = torch.zeros(min(m, n), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        h.geqrf(A, tau)
        h.synchronize()
        return A.cpu().numpy(), tau.cpu().numpy()
    finally:
        h.close()


def check(a, af, tau, tag):
    m, n = a.shape
    k = min(m, n)
    af_o, tau_o = O.factor(a)
    e_r = rel_err_colscaled(np.triu(af[:k]), np.triu(af_o[:k]))
    e_t = float(np.max(np.abs(tau - tau_o)))
    q = R.generateq(af[:, :k], tau, n=k)
    res, orth = qr_residual(a, q, np.triu(af[:k])), orth_error(q)
    print(f"  {tag}: R {e_r:.1e} tau {e_t:.1e} residual {res:.2f} orth {orth:.2f}", flush=True)
    assert e_r <= 1e-11 and e_t <= 1e-11 and res <= 10 and orth <= 10


@pytest.mark.parametrize("m,n", [(300, 64), (1000, 200), (2048, 2048), (4097, 96), (600, 900)])
def test_accepted_panels(m, n):
    a = cmat(m, n, m + n)
    print(f"[case] accepted {m}x{n}", flush=True)
    af, tau = factor_dev(a)
    check(a, af, tau, "cholqr2")
    af1, tau1 = factor_dev(a, panel=1)  # column by column only
    assert np.abs(af - af1).max() <= 1e-11 * np.abs(af1).max()


@pytest.mark.parametrize("kind", ["tiny", "zero", "repeated", "huge"])
def test_rejected_panels(kind):
    m, n = 800, 160
    a = cmat(m, n, 7)
    if kind == "tiny":
        a[:, 37] *= 1e-20
    elif kind == "zero":
        a[:, 70] = 0
    elif kind == "repeated":
        a[:, 101] = a[:, 99]
    else:
        a[:, 5] *= 1e200
    print(f"[case] rejected ({kind}) {m}x{n}", flush=True)
    af, tau = factor_dev(a)
    if kind == "repeated":
        # rank deficient: compare up to the repeated column, then the identities
        af_o, tau_o = O.factor(a)
        assert rel_err_colscaled(np.triu(af[:100, :100]), np.triu(af_o[:100, :100])) <= 1e-11
        q = R.generateq(af[:, :n], tau, n=n)
        assert qr_residual(a, q, np.triu(af[:n])) <= 10 and orth_error(q) <= 10
    else:
        check(a, af, tau, kind)
