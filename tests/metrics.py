"""Error metrics shared by the parity tests (the north-star criteria).

  factorization residual  ||A - Q R||_F / (n * eps * ||A||_F)      <= 10
  orthogonality           ||Q^H Q - I||_F / (n * eps)               <= 10
  elementwise relative error of R and tau against the oracle, reported
  relative to the column scale (max |R(:,j)|), since entries that cancel to
  tiny values have no relative accuracy in any backward-stable algorithm.
"""
import numpy as np

EPS = np.finfo(np.float64).eps


def _snorm(x):
    """Frobenius norm computed on a power-of-two-scaled copy (no overflow)."""
    mx = float(np.max(np.abs(x))) if x.size else 0.0
    if mx == 0.0 or not np.isfinite(mx):
        return mx
    e = np.frexp(mx)[1]
    return float(np.ldexp(np.linalg.norm(np.ldexp(x, -e) if x.dtype.kind == "f" else x * 2.0 ** -e), e))


def qr_residual(a, q, r):
    n = max(a.shape[1], 1)
    na = _snorm(a)
    if na == 0:
        return _snorm(a - q @ r)
    mx = float(np.max(np.abs(a)))
    e = np.frexp(mx)[1]
    s = 2.0 ** -e
    # scale A by a power of two before the product so Q R cannot overflow
    d = (a * s) - q @ (r * s)
    return float(_snorm(d) / (n * EPS * _snorm(a * s)))


def orth_error(q):
    n = max(q.shape[1], 1)
    return float(np.linalg.norm(q.conj().T @ q - np.eye(q.shape[1])) / (n * EPS))


def rel_err_colscaled(x, ref):
    """max_ij |x_ij - ref_ij| / max_i |ref_ij| (column-scaled elementwise error)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    if ref.ndim == 1:
        s = np.max(np.abs(ref)) if ref.size else 1.0
        return float(np.max(np.abs(x - ref)) / (s if s > 0 else 1.0)) if ref.size else 0.0
    s = np.max(np.abs(ref), axis=0)
    s[s == 0] = 1.0
    return float(np.max(np.abs(x - ref) / s[None, :])) if ref.size else 0.0


def rel_err_elem(x, ref, floor=1e-300):
    x = np.asarray(x)
    ref = np.asarray(ref)
    d = np.abs(x - ref) / np.maximum(np.abs(ref), floor)
    return float(np.max(d)) if d.size else 0.0


def r_of(afac):
    m, n = afac.shape
    k = min(m, n)
    return np.triu(afac[:k, :])
