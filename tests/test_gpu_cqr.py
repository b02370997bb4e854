"""The CholeskyQR2 + Householder-reconstruction panel against the oracle.

Real 64-column panels take the CholeskyQR2 kernel by default; tuning
panel=1 forces the column-by-column kernel.  Both must match the oracle's
column-by-column reflectors (R, V, tau elementwise, column-scaled, <= 1e-11)
and each other; panels the kernel rejects (singular or badly conditioned
Gram, extreme scaling) must come out of the column-wise fallback with the
same parity.  Results are deterministic run to run.
"""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu

TOL_ELEM = 1e-11


def _factor_dev(a, panel):
    """panel 1: column-wise; 2: CholeskyQR2 per 64-column panel (64-column
    blocks); 0: the default, 128-column outer blocks in one launch each."""
    import torch
    m, n = a.shape
    h = R.Handle()
    h.set_tuning(nb=0 if panel == 0 else 64, panel=panel)
    d = R.colmajor_empty(m, n, torch.float64)
    d.copy_(torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda").t())
    tau = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    h.geqrf(d, tau)
    h.synchronize()
    return d.cpu().numpy(), tau.cpu().numpy()


def _compare(a, af, tau, tag):
    m, n = a.shape
    k = min(m, n)
    af_o, tau_o = O.factor(a)
    e_r = rel_err_colscaled(np.triu(af[:k]), np.triu(af_o[:k]))
    e_t = float(np.max(np.abs(tau - tau_o)))
    e_v = float(np.max(np.abs(np.tril(af, -1) - np.tril(af_o, -1))))
    q = R.generateq(af, tau, n=k) if m >= n else None
    out = {"R": e_r, "tau": e_t, "V": e_v}
    if q is not None:
        out["res"] = qr_residual(a, q, np.triu(af[:k]))
        out["orth"] = orth_error(q)
    print(f"        {tag}: {out}", flush=True)
    assert e_r <= TOL_ELEM and e_t <= TOL_ELEM and e_v <= 1e-9, out
    if q is not None:
        assert out["res"] <= 10 and out["orth"] <= 10, out


@pytest.mark.parametrize("shape", [(128, 64), (200, 64), (512, 256), (1000, 300), (4096, 640), (20000, 192)])
def test_cqr_panel_matches_oracle(shape):
    m, n = shape
    print(f"[case] cqr {shape}", flush=True)
    iseed = np.array([m % 4096, n % 4096, 7, 9], dtype=np.int32)
    a = O.generate_matrix(O.NORMAL01, m, n, iseed)
    af2, tau2 = _factor_dev(a, panel=2)
    _compare(a, af2, tau2, "cholqr2")
    af1, tau1 = _factor_dev(a, panel=1)
    _compare(a, af1, tau1, "column-wise")
    k = min(m, n)
    assert rel_err_colscaled(np.triu(af2[:k]), np.triu(af1[:k])) <= TOL_ELEM
    af3, tau3 = _factor_dev(a, panel=2)
    assert np.array_equal(af3, af2) and np.array_equal(tau3, tau2), "not deterministic"
    af0, tau0 = _factor_dev(a, panel=0)
    _compare(a, af0, tau0, "outer-block kernel")
    assert rel_err_colscaled(np.triu(af0[:k]), np.triu(af1[:k])) <= TOL_ELEM
    af4, tau4 = _factor_dev(a, panel=0)
    assert np.array_equal(af4, af0) and np.array_equal(tau4, tau0), "outer-block kernel not deterministic"


@pytest.mark.parametrize("panel", [2, 0])
def test_cqr_sign_rule_zero_pivot(panel):
    """A panel whose running pivot is exactly zero (alpha == 0): D = +1, the
    reference's beta > 0."""
    print(f"[case] cqr zero pivot panel={panel}", flush=True)
    rng = np.random.default_rng(3)
    a = np.asfortranarray(rng.standard_normal((256, 128)))
    a[0, 0] = 0.0
    af, tau = _factor_dev(a, panel=panel)
    _compare(a, af, tau, "zero pivot")
    assert af[0, 0] > 0


@pytest.mark.parametrize("panel", [2, 0])
@pytest.mark.parametrize("kind", ["ill", "dependent", "scaled", "second"])
def test_cqr_rejected_panels_fall_back(kind, panel):
    """Badly conditioned, rank-deficient and extremely scaled panels go to the
    column-wise kernel and still match the oracle."""
    print(f"[case] cqr fallback {kind} panel={panel}", flush=True)
    rng = np.random.default_rng(11)
    m, n = 600, 192
    a = rng.standard_normal((m, n))
    if kind == "ill":
        a[:, 70] = a[:, 69] + 1e-9 * a[:, 70]          # second panel: kappa ~ 1e9
    elif kind == "dependent":
        a[:, 5] = 2.0 * a[:, 2]
    elif kind == "second":
        a[:, 100] = a[:, 99] - 3.0 * a[:, 98]           # second panel of the first block
    else:
        a[:, 130] *= 1e200                              # third panel: Gram would overflow
    a = np.asfortranarray(a)
    af, tau = _factor_dev(a, panel=panel)
    if kind in ("dependent", "ill", "second"):
        # the reflectors after a (near) dependence are fixed by rounding noise
        # in any implementation: elementwise before it, normwise overall
        c = {"dependent": 5, "ill": 69, "second": 99}[kind]
        af_o, tau_o = O.factor(a)
        assert rel_err_colscaled(np.triu(af[:n])[:, :c], np.triu(af_o[:n])[:, :c]) <= TOL_ELEM
        q = R.generateq(af, tau, n=n)
        assert qr_residual(a, q, np.triu(af[:n])) <= 10 and orth_error(q) <= 10
    else:
        _compare(a, af, tau, kind)
