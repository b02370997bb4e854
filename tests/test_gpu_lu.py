"""GPU parity of LU::Factor and LU::Solve (RNP/LA/LU.hpp:77-144) against the
oracle's restatement on the same seeded inputs.

Bar: the same pivots (partial pivoting picks the same rows), the packed L and
U within 1e-11 of the largest entry of A (times max(m,n)/100 for growth of
the rounding over the panels), the reference's return value exactly, and
op(A) X = B solutions within 1e-11 relative of the oracle's.
"""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R

pytestmark = pytest.mark.gpu


def seeded(m, n, cplx, seed):
    iseed = np.array([seed % 4096, m % 4096, n % 4096, 9], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed, complex_=cplx)


def check_lu(a, tag):
    print(f"[case] {tag} shape={a.shape} dtype={a.dtype}", flush=True)
    m, n = a.shape
    lo, po, io = O.lu_factor(a)
    lg, pg, ig = R.lu_factor(a)
    scale = max(np.abs(a).max(), 1e-300) * max(1.0, max(m, n) / 100.0)
    e = float(np.abs(lg - lo).max() / scale) if lo.size else 0.0
    same = np.array_equal(po, pg)
    print(f"        pivots equal {same}, |LU - LU_o| {e:.2e}, info {ig} (oracle {io})", flush=True)
    assert same and e <= 1e-11 and ig == io
    return lg, pg, ig


SHAPES = [(5, 5), (64, 64), (65, 65), (300, 200), (200, 300), (1, 1), (1, 9), (9, 1), (1000, 1000), (2048, 2048),
          (4096, 300), (513, 1200)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_lu_factor_matches_oracle(shape, cplx):
    check_lu(seeded(*shape, cplx, 17), f"lu cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
def test_lu_zero_pivots_and_return_value(cplx):
    """Exact zero columns: the reference's return arithmetic (0 for a zero
    pivot in column 0 only, z+1 unblocked, iinfo + j + 1 blocked)."""
    a = seeded(20, 20, cplx, 3)
    a[:, 0] = 0.0
    check_lu(a, f"lu zero col 0 cplx={cplx}")
    a = seeded(20, 20, cplx, 4)
    a[:, 6] = 0.0
    a[:, 9] = 0.0
    check_lu(a, f"lu zero cols 6, 9 cplx={cplx}")
    a = seeded(200, 200, cplx, 5)
    a[:, 70] = 0.0
    a[:, 150] = 0.0
    check_lu(a, f"lu zero cols 70, 150 cplx={cplx}")
    a = np.zeros((130, 130), dtype=np.complex128 if cplx else np.float64)
    check_lu(a, f"lu zero matrix cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_lu_solve_matches_oracle(trans, cplx):
    print(f"[case] lu solve {trans} cplx={cplx}", flush=True)
    rng = np.random.default_rng(11)
    for n, nrhs in ((40, 3), (700, 5)):
        a = seeded(n, n, cplx, n)
        b = rng.standard_normal((n, nrhs))
        if cplx:
            b = b + 1j * rng.standard_normal((n, nrhs))
        lo, po, _ = O.lu_factor(a)
        want = O.lu_solve(trans, lo, po, b)
        got = R.lu_solve(trans, lo, po, b)
        err = float(np.max(np.abs(got - want)) / max(np.abs(want).max(), 1e-300))
        print(f"        n={n}: {err:.2e}", flush=True)
        assert err <= 1e-11


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_lu_refine_matches_oracle(trans, cplx):
    """RefineSolution from the same starting x: the refined x within 1e-12
    of the oracle's, berr at working precision in both, ferr within 5% (its
    weights |r| + (n+1) eps (|op(A)||x| + |b|) carry the final residual r,
    which is rounding noise of relative size up to ~1/(n+1))."""
    print(f"[case] lu refine {trans} cplx={cplx}", flush=True)
    rng = np.random.default_rng(21)
    for n in (30, 500):
        a = seeded(n, n, cplx, n + 1)
        b = rng.standard_normal((n, 2))
        if cplx:
            b = b + 1j * rng.standard_normal((n, 2))
        lo, po, _ = O.lu_factor(a)
        x0 = O.lu_solve(trans, lo, po, b)
        # start from a perturbed solution so that refinement has work to do
        x0 = x0 * (1.0 + 1e-9 * rng.standard_normal(x0.shape))
        xo, fo, bo = O.lu_refine(trans, a, lo, po, b, x0.copy())
        xg, fg, bg = R.lu_refine(trans, a, lo, po, b, x0.copy())
        ex = float(np.max(np.abs(xg - xo)) / np.abs(xo).max())
        ef = float(np.max(np.abs(fg - fo) / fo))
        print(f"        n={n}: x {ex:.2e} ferr {ef:.2e} berr {bg} (oracle {bo})", flush=True)
        assert ex <= 1e-12 and ef <= 0.05
        assert np.all(bg <= 4 * np.finfo(float).eps) and np.all(bo <= 4 * np.finfo(float).eps)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 5, 64, 65, 300, 1000, 2048])
def test_lu_invert_matches_oracle(n, cplx):
    """LU::Invert (LU.hpp:146-214) through the drop-in ABI against the
    oracle's restatement of the same factors: entries within 1e-11 times
    kappa-scaled rounding of the largest entry, and inv(A) A = I."""
    a = seeded(n, n, cplx, 31 + n)
    print(f"[case] LU::Invert n={n} cplx={cplx}", flush=True)
    lu, piv, info = O.lu_factor(a)
    assert info == 0
    inv_o, st_o = O.lu_invert(lu, piv)
    inv_g, st_g = R.lu_invert(lu, piv)
    scale = np.abs(inv_o).max()
    e = float(np.abs(inv_g - inv_o).max() / scale)
    resid = float(np.abs(inv_g @ a - np.eye(n)).max())
    kappa = float(np.linalg.cond(a)) if n <= 1000 else 1e4
    print(f"        status {st_g} (oracle {st_o}), |inv - inv_o| / |inv_o| {e:.2e}, |inv A - I| {resid:.2e}, "
          f"kappa {kappa:.1e}", flush=True)
    assert st_g == st_o == 0
    assert e <= 1e-13 * n * kappa and resid <= 1e-13 * n * kappa


@pytest.mark.parametrize("cplx", [False, True])
def test_lu_invert_zero_pivot(cplx):
    """A zero U_ii returns -(i+1) and leaves the factors as they were, as
    Triangular::Invert does (Triangular.hpp:217-223)."""
    print(f"[case] LU::Invert zero pivot cplx={cplx}", flush=True)
    a = seeded(50, 50, cplx, 5)
    lu, piv, _ = O.lu_factor(a)
    lu[17, 17] = 0.0
    out, st = R.lu_invert(lu, piv)
    assert st == -18 and np.array_equal(out, lu)
