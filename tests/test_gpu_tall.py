"""Factor of matrices taller than a 64-column panel holds (more than 32768
rows): the drop-in C ABI (rnpx_dgeqrf / rnpx_zgeqrf on host memory) and the
device handle, against the oracle on the leading columns (column-causal) and
through ||A - QR|| / (n eps ||A||) <= 10 and ||Q^H Q - I|| / (n eps) <= 10 with
Q from GenerateQ.  The reference's QR::Factor (RNP/LA/QR.hpp:129-197) has no
row limit.
"""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu


def seeded(m, n, cplx, seed):
    iseed = np.array([seed % 4096, (seed * 7) % 4096, (seed * 13) % 4096, (2 * seed + 1) % 4096], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed, cplx)


@pytest.mark.parametrize("m,n,cplx", [(65536, 64, False), (100000, 128, False), (40000, 200, False),
                                      (70000, 40, True)])
def test_tall_factor_drop_in(m, n, cplx):
    print(f"[case] tall drop-in Factor {m}x{n} cplx={cplx}", flush=True)
    a = seeded(m, n, cplx, m + n)
    af, tau = R.factor(a)                     # rnpx_{d,z}geqrf through the C ABI
    lead = min(n, 64)
    af_o, tau_o = O.factor(np.asfortranarray(a[:, :lead]))
    e_r = rel_err_colscaled(np.triu(af[:lead, :lead]), np.triu(af_o[:lead]))
    e_v = float(np.max(np.abs(np.tril(af[:, :lead], -1) - np.tril(af_o, -1))))
    e_t = float(np.max(np.abs(tau[:lead] - tau_o)))
    q = R.generateq(af, tau, n=n)
    res, orth = qr_residual(a, q, np.triu(af[:n])), orth_error(q)
    out = {"R_lead": e_r, "V_lead": e_v, "tau_lead": e_t, "res": res, "orth": orth}
    print(f"        {out}", flush=True)
    assert e_r <= 1e-11 and e_v <= 1e-11 and e_t <= 1e-11, out
    assert res <= 10 and orth <= 10, out


def test_tall_least_squares():
    """The least-squares driver on a 120000 x 96 problem (Q^H b formed inside
    the factorization) against numpy's solution."""
    import torch
    m, n = 120000, 96
    print(f"[case] tall LS {m}x{n}", flush=True)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((m, n))
    b = a @ rng.standard_normal((n, 2)) + 1e-3 * rng.standard_normal((m, 2))
    A = R.colmajor_empty(m, n, torch.float64)
    A.copy_(torch.from_numpy(a))
    B = R.colmajor_empty(m, 2, torch.float64)
    B.copy_(torch.from_numpy(b))
    tau = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    h = R.Handle()
    h.gels(A, tau, B)
    h.synchronize()
    x = B[:n].cpu().numpy()
    x_ref = np.linalg.lstsq(a, b, rcond=None)[0]
    err = float(np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)))
    print(f"        |x - x_numpy| / |x| = {err:.2e}", flush=True)
    assert err <= 1e-12
