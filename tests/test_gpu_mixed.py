"""GPU tests of the mixed-precision least-squares form (FP32-MFMA trailing
updates, FP64 refinement) against the oracle's FP64 solution."""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R

pytestmark = pytest.mark.gpu


def conditioned(m, n, cond, seed):
    """A = U diag(s) V^T with s log-spaced from 1 to 1/cond and U, V the Q
    factors of Gaussian matrices (the mixed least-squares configuration)."""
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((m, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(cond), n)
    return (u * s) @ v.T


def berr(a, b, x):
    """eta = ||A^T r|| / (||A||_F (||A||_F ||x|| + ||b||)), worst column."""
    r = b - a @ x
    an = np.linalg.norm(a)
    return float(np.max(np.linalg.norm(a.T @ r, axis=0) / (an * (an * np.linalg.norm(x, axis=0) + np.linalg.norm(b, axis=0)))))


def berr_ls(a, b, x):
    """||A^T r|| / (||A||_F ||r||): the stricter measure when r != 0."""
    r = b - a @ x
    return float(np.max(np.linalg.norm(a.T @ r, axis=0) / (np.linalg.norm(a) * np.linalg.norm(r, axis=0))))


def dev(x):
    import torch
    t = R.colmajor_empty(x.shape[0], x.shape[1], torch.float64)
    t.copy_(torch.from_numpy(np.ascontiguousarray(x)))
    torch.cuda.synchronize()  # the handle's stream reads it next
    return t


@pytest.mark.parametrize("cond", [1e2, 1e6])
@pytest.mark.parametrize("m,n,nrhs", [(2048, 256, 1), (3000, 700, 4), (2500, 1000, 6)])
def test_mixed_ls_converges(m, n, nrhs, cond):
    import torch
    print(f"[case] mixed LS {m}x{n} nrhs={nrhs} cond={cond:g}", flush=True)
    a = conditioned(m, n, cond, 1)
    rng = np.random.default_rng(2)
    xt = rng.standard_normal((n, nrhs))
    b = a @ xt + 1e-8 * rng.standard_normal((m, nrhs))
    A, B = dev(a), dev(b)
    X = R.colmajor_empty(n, nrhs, torch.float64)
    h = R.Handle()
    st, iters, be = h.gels_mixed(A, B, X, max_iter=40, tol=1e-14)
    h.synchronize()
    x = X.cpu().numpy()
    be_host = berr(a, b, x)
    x_o = O.ls_solve(a, b)
    fwd = float(np.max(np.abs(x - x_o)) / np.max(np.abs(x_o)))
    ls_g, ls_o = berr_ls(a, b, x), berr_ls(a, b, x_o)
    print(f"        iters={iters} eta(dev)={be:.2e} eta(host)={be_host:.2e} eta_ls gpu={ls_g:.2e} oracle={ls_o:.2e} "
          f"|x-x_oracle|/|x|={fwd:.2e}", flush=True)
    assert be <= 1e-14 and be_host <= 1e-13
    # each step contracts the error by about kappa * 2^-24 (the FP32 updates;
    # R^{-1} applied exactly in FP64): a wrong block of R^{-1} shows as many
    # more steps
    assert iters <= (4 if cond < 1e3 else 12)
    # refined to working precision: as backward-stable as the FP64 oracle
    assert ls_g <= max(10 * ls_o, 1e-13)
    # inputs untouched
    assert np.array_equal(A.cpu().numpy(), a) and np.array_equal(B.cpu().numpy(), b)


def test_ls_berr_matches_host():
    import torch
    a = conditioned(500, 100, 1e3, 3)
    b = np.random.default_rng(4).standard_normal((500, 2))
    x = np.linalg.lstsq(a, b, rcond=None)[0] + 1e-9
    h = R.Handle()
    got = h.ls_berr(dev(a), dev(b), dev(x))
    assert got == pytest.approx(berr(a, b, x), rel=1e-6)
