"""Matrices holding Inf or NaN where the reference's zero trims decide the
result: dense leading rows, zero rows below them except a few non-finite
entries (so the reflectors of the columns before those entries end in zeros
and skip the rows that hold them), sometimes zero trailing columns.  Shared
by the oracle-vs-reference check and the device parity test."""
import numpy as np

INF, NAN = np.inf, np.nan

# (tag, m, n, dense rows k0, [(row, col, value)], zero trailing columns, complex, lwork per column or 0)
FIXED = [
    ("inf at a panel start", 300, 200, 180, [(250, 128, INF)], 0, False, 0),
    ("inf mid panel", 300, 200, 180, [(250, 40, INF)], 0, False, 0),
    ("nan", 300, 200, 180, [(250, 64, NAN)], 0, False, 0),
    ("inf in the unblocked tail", 300, 260, 220, [(290, 150, INF)], 0, False, 0),
    ("all unblocked (k <= 128)", 100, 80, 60, [(90, 10, INF)], 0, False, 0),
    ("non-finite x of the first reflector", 200, 150, 150, [(180, 0, NAN)], 0, False, 0),
    ("zero trailing columns", 220, 170, 220, [(5, 0, NAN)], 60, False, 0),
    ("wide (m < n)", 150, 300, 100, [(120, 20, INF)], 0, False, 0),
    ("short workspace (nb 16)", 300, 200, 180, [(250, 128, INF)], 0, False, 16),
    ("-inf and nan", 320, 240, 200, [(300, 90, -INF), (260, 170, NAN)], 20, False, 0),
    ("complex inf", 260, 180, 150, [(200, 70, complex(INF, 0.0))], 0, True, 0),
    ("complex nan", 240, 170, 130, [(180, 33, complex(0.0, NAN))], 15, True, 0),
]


def build(m, n, k0, specials, zero_cols, cplx, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m, n))
    if cplx:
        a = a + 1j * rng.standard_normal((m, n))
    a[k0:, :] = 0
    if zero_cols:
        a[:, n - zero_cols:] = 0
    for r, c, v in specials:
        a[r, c] = v
    return np.asfortranarray(a)


def random_cases(count, seed=11):
    """Random shapes with one to three non-finite entries in the zero rows."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        m = int(rng.integers(150, 400))
        n = int(rng.integers(100, m))
        k0 = int(rng.integers(n // 2, m - 5))
        specials = [(int(rng.integers(k0, m)), int(rng.integers(0, n)), [INF, -INF, NAN][int(rng.integers(0, 3))])
                    for _ in range(int(rng.integers(1, 4)))]
        out.append((f"random {t}", m, n, k0, specials, 10 if t % 3 == 0 else 0, False, 0))
    return out


def cases():
    return [(c, i) for i, c in enumerate(FIXED + random_cases(8))]


def compare(af, tau, af_ref, tau_ref, tol=1e-11):
    """Same entries non-finite, finite entries equal to tol (relative to the
    largest finite reference entry).  Returns the error."""
    fin, fin_ref = np.isfinite(af), np.isfinite(af_ref)
    assert (fin == fin_ref).all(), f"{int((fin != fin_ref).sum())} entries differ in finiteness"
    assert (np.isfinite(tau) == np.isfinite(tau_ref)).all(), "tau differs in finiteness"
    scale = max(1.0, float(np.abs(af_ref[fin_ref]).max(initial=0.0)))
    err = float(np.abs(af[fin] - af_ref[fin]).max(initial=0.0)) / scale
    tf = np.isfinite(tau_ref)
    err = max(err, float(np.abs(tau[tf] - tau_ref[tf]).max(initial=0.0)))
    assert err <= tol, err
    return err
