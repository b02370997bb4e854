"""Reflector::LastNonzero* (RNP/LA/Reflector.hpp:36-75) on the device, and
QR::Factor of matrices holding Inf or NaN, where the reference's zero trims
(Apply 196-221, GenerateBlockTr 581-623, ApplyBlock 742-744) and its blocking
(nb 32, crossover 128) decide which 0 * Inf products are formed.

Limits: bit-exact against the oracle.  Factor: the same entries non-finite
as the oracle's, finite entries within 1e-11 (relative to the largest), tau
likewise; the oracle's handling of these matrices is itself pinned against
the reference built from its sources (tests/test_oracle_refqr.py)."""
import numpy as np
import pytest

import nonfinite_cases as NF
import oracle as O
import rnp2_amd as R

pytestmark = pytest.mark.gpu


def limit_patterns():
    rng = np.random.default_rng(3)
    out = []
    for t in range(40):
        m, n = int(rng.integers(1, 70)), int(rng.integers(1, 70))
        a = np.zeros((m, n))
        kind = t % 5
        if kind == 0:  # a few scattered nonzeros
            for _ in range(int(rng.integers(1, 4))):
                a[int(rng.integers(0, m)), int(rng.integers(0, n))] = rng.standard_normal()
        elif kind == 1:  # dense block in the top-left
            a[:int(rng.integers(1, m + 1)), :int(rng.integers(1, n + 1))] = rng.standard_normal()
        elif kind == 2:  # NaN alone
            a[int(rng.integers(0, m)), int(rng.integers(0, n))] = np.nan
        elif kind == 3:  # negative zeros only
            a[:, :] = -0.0
        else:  # corners set
            a[m - 1, n - 1] = 1.0 if t % 2 else 0.0
            a[0, n - 1] = 0.0
            a[m - 1, 0] = 2.0
        out.append(a)
    out.append(np.zeros((5, 7)))
    return out


@pytest.mark.parametrize("cplx", [False, True])
def test_last_nonzero_limits_match_oracle(cplx):
    for i, a in enumerate(limit_patterns()):
        if cplx:
            a = a.astype(complex)
            if i % 7 == 3:
                a[np.nonzero(a)] *= 1j  # purely imaginary entries count as nonzero
        print(f"[case] limits {i} shape={a.shape} cplx={cplx}", flush=True)
        want = (O.last_nonzero_column_limit(a), O.last_nonzero_row_limit(a))
        got = R.last_nonzero_limits(a)
        assert got == want, (i, got, want)


@pytest.mark.parametrize("case,seed", NF.cases(), ids=lambda c: c[0] if isinstance(c, tuple) else str(c))
def test_factor_nonfinite_matches_oracle(case, seed):
    tag, m, n, k0, specials, zc, cplx, lwn = case
    a = NF.build(m, n, k0, specials, zc, cplx, seed)
    print(f"[case] {tag} shape={a.shape} cplx={cplx} lwork/n={lwn}", flush=True)
    before = R.exact_path_count()
    if lwn:
        af_g = a.copy(order="F")
        tau_g = np.zeros(min(m, n), dtype=a.dtype)
        lw = np.array([lwn * n], dtype=np.uint64)
        work = np.zeros(lwn * n, dtype=a.dtype)
        assert R.Factor(m, n, af_g, m, tau_g, lw, work) == 0
        af_o, tau_o = O.factor(a, lwork=lwn * n)
    else:
        af_g, tau_g = R.factor(a)
        af_o, tau_o = O.factor(a)
    assert R.exact_path_count() == before + 1
    err = NF.compare(af_g, tau_g, af_o, tau_o)
    print(f"  finite entries: max error {err:.2e}; non-finite {int((~np.isfinite(af_o)).sum())}", flush=True)


def test_finite_matrix_takes_the_fast_path():
    rng = np.random.default_rng(0)
    a = np.asfortranarray(rng.standard_normal((400, 300)))
    a[350:, :] = 0.0  # zero tails alone change nothing the trims would skip
    print("[case] finite with zero tail 400x300", flush=True)
    before = R.exact_path_count()
    af_g, tau_g = R.factor(a)
    assert R.exact_path_count() == before
    af_o, tau_o = O.factor(a)
    assert np.isfinite(af_g).all()
    k = 300
    err = np.abs(np.triu(af_g[:k]) - np.triu(af_o[:k])).max() / np.abs(af_o).max()
    assert err <= 1e-11, err
