"""Outer block widths 32, 64, 128 and 256 (rnpx_tuning.nb): the blocked
Factor (QR.hpp:129-197) against the oracle for each, including heights and
widths that leave a partial last block, complex matrices and look-ahead off.
R and tau column-scaled elementwise error <= 1e-11, ||A - QR|| / (n eps ||A||)
<= 10 and ||Q^H Q - I|| / (n eps) <= 10."""
import numpy as np
import pytest
import torch

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu

SHAPES = [(1000, 700), (777, 777), (2048, 1300), (600, 900)]


def seeded(m, n, cplx, seed):
    iseed = np.array([seed % 4096, (seed * 7) % 4096, (seed * 13) % 4096, (2 * seed + 1) % 4096], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed, cplx)


@pytest.mark.parametrize("nb", [32, 64, 128, 256])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("lookahead", [0, -1])
def test_block_width(nb, cplx, lookahead):
    h = R.Handle()
    try:
        h.set_tuning(nb=nb, lookahead=lookahead)
        for s, (m, n) in enumerate(SHAPES):
            a = seeded(m, n, cplx, 31 * s + nb)
            print(f"[case] nb={nb} lookahead={lookahead} {m}x{n} cplx={cplx}", flush=True)
            dt = torch.complex128 if cplx else torch.float64
            A = R.colmajor_empty(m, n, dt)
            A.copy_(torch.from_numpy(np.ascontiguousarray(a)).to("cuda"))
            torch.cuda.synchronize()  # the copy runs on torch's stream, geqrf on the handle's
            k = min(m, n)
            tau = torch.empty(k, dtype=dt, device="cuda")
            h.geqrf(A, tau)
            h.synchronize()
            af = A.cpu().numpy()
            tg = tau.cpu().numpy()
            af_o, tau_o = O.factor(a)
            assert rel_err_colscaled(np.triu(af[:k]), np.triu(af_o[:k])) <= 1e-11
            assert np.max(np.abs(tg - tau_o)) <= 1e-11
            q = R.generateq(af[:, :k], tg, n=k)
            assert qr_residual(a, q, np.triu(af[:k])) <= 10
            assert orth_error(q) <= 10
    finally:
        h.close()
