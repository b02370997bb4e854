"""The oracle's QR restatement pinned against the real reference's blocked QR.

Two sources, both the reference's own RNP/LA/QR.hpp built with its
src/BLAS_mix.cpp against the image's MKL:
  * tests/golden/rnp_reference_qr.npz (written by make_golden_qr.py), which
    always runs: config 1 (512 x 256, RNP stream), the reference tests'
    shapes, edge matrices, MultQ's four products, a short workspace and
    problems drawn from the batched configuration;
  * the live library oracle/_ref/librnpref_qr.so on further shapes, where it
    was built and MKL loads (skipped otherwise).
The two differ from the oracle only in the Level-2/3 BLAS summation order
(MKL against the oracle's own loops), so they agree to a few ulps of the
column scale.  Also here: the CPU-baseline routine factor_panels pinned
bit for bit against Factor, and RandomMatrix against the reference's.
"""
import os

import numpy as np
import pytest

import oracle as O
from oracle import refqr
from metrics import rel_err_colscaled

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "rnp_reference_qr.npz"))
TOL = 1e-13


def rnp_stream(m, n, cplx):
    O.reset_internal_seed()
    return O.generate_matrix(O.UNIFORM_11, m, n, None, cplx)


def seeded(m, n, p):
    iseed = np.array([p % 4096, (7 * p + 1) % 4096, (13 * p + 2) % 4096, (2 * p + 1) % 4096], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed)


def factor_errors(af, tau, af_ref, tau_ref):
    k = min(af.shape)
    e_r = rel_err_colscaled(np.triu(af[:k]), np.triu(af_ref[:k]))
    e_v = float(np.max(np.abs(np.tril(af, -1) - np.tril(af_ref, -1)))) if k else 0.0
    e_t = float(np.max(np.abs(tau - tau_ref))) if k else 0.0
    return e_r, e_v, e_t


def test_config1_against_reference_fixture():
    a = rnp_stream(512, 256, False)
    af, tau = O.factor(a)
    e = factor_errors(af, tau, G["c1_afac"], G["c1_tau"])
    q = O.generateq(af, tau, n=256)
    e_q = float(np.max(np.abs(q - G["c1_q"])))
    print("c1", e, e_q)
    assert max(e) <= TOL and e_q <= TOL


@pytest.mark.parametrize("key,m,n,cplx", [("reft_d_170x160", 170, 160, False), ("reft_d_166x170", 166, 170, False),
                                          ("reft_z_170x160", 170, 160, True)])
def test_reference_test_shapes_against_fixture(key, m, n, cplx):
    af, tau = O.factor(rnp_stream(m, n, cplx))
    e = factor_errors(af, tau, G[key + "_afac"], G[key + "_tau"])
    print(key, e)
    assert max(e) <= TOL


def numerical_rank(af):
    """k minus the number of trailing diagonal entries of R at the noise
    level (a zero column inside the matrix does not count)."""
    k = min(af.shape)
    d = np.abs(np.diag(af[:k, :k]))
    tiny = d <= 1e-10 * max(np.max(d), 1e-300) if k else d
    r = k
    while r > 0 and tiny[r - 1]:
        r -= 1
    return r


def compare_edge(a, af, tau, af_r, tau_r):
    """Factor errors on a possibly rank-deficient matrix: past the numerical
    rank r the reflectors are those of rounding noise and depend on the
    summation order, so only R(:r, :), V(:, :r) and tau(:r) are compared, and
    R(r:, r:) must be at the noise level in both."""
    k = min(a.shape)
    r = numerical_rank(af_r)
    if r == k:
        return factor_errors(af, tau, af_r, tau_r)
    e_r = rel_err_colscaled(np.triu(af[:r]), np.triu(af_r[:r]))
    e_v = float(np.max(np.abs(np.tril(af, -1)[:, :r] - np.tril(af_r, -1)[:, :r])))
    e_t = float(np.max(np.abs(tau[:r] - tau_r[:r])))
    scale = np.max(np.abs(a))
    noise = max(np.max(np.abs(np.triu(af[r:k, r:]))), np.max(np.abs(np.triu(af_r[r:k, r:]))))
    assert noise <= 1e-12 * scale, noise
    return e_r, e_v, e_t


@pytest.mark.parametrize("pre", ["d", "z"])
def test_edge_matrices_against_fixture(pre):
    for i in range(int(G[f"edge_{pre}_count"])):
        a = G[f"edge_{pre}_{i}_a"]
        af, tau = O.factor(a)
        af_r, tau_r = G[f"edge_{pre}_{i}_afac"], G[f"edge_{pre}_{i}_tau"]
        e = compare_edge(a, af, tau, af_r, tau_r)
        r = numerical_rank(af_r)
        if r < min(a.shape):
            print(pre, i, a.shape, "rank", r, e)
            assert max(e) <= TOL, (i, e)
            continue
        print(pre, i, a.shape, e)
        # V entries scale with 1/(alpha - beta): compare them relative to the column of V
        k = min(a.shape)
        vs = np.max(np.abs(np.tril(af_r, -1)), axis=0)[:k] if a.shape[0] > 1 else np.zeros(k)
        vs[vs == 0] = 1.0
        e_v = float(np.max(np.abs(np.tril(af - af_r, -1))[:, :k] / vs)) if a.shape[0] > 1 else 0.0
        assert e[0] <= TOL and e[2] <= TOL and e_v <= 1e-12, (i, e, e_v)
        # the sign rule and tau = 0 decisions are exact
        assert np.array_equal(tau == 0, tau_r == 0)
        d, d_r = np.diag(af[:k, :k]), np.diag(af_r[:k, :k])
        assert np.array_equal(np.sign(d.real), np.sign(d_r.real))


@pytest.mark.parametrize("pre", ["d", "z"])
def test_multq_against_fixture(pre):
    af, tau = O.factor(G[f"mq_{pre}_a"])
    assert max(factor_errors(af, tau, G[f"mq_{pre}_afac"], G[f"mq_{pre}_tau"])) <= TOL
    for side in "LR":
        for trans in "NC":
            c = G[f"mq_{pre}_{side}{trans}_c"]
            out = O.multq(side, trans, G[f"mq_{pre}_afac"], G[f"mq_{pre}_tau"], c)
            ref = G[f"mq_{pre}_{side}{trans}_out"]
            e = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
            print(pre, side, trans, e)
            assert e <= TOL


def test_short_workspace_against_fixture():
    a = seeded(200, 140, 999)
    af, tau = O.factor(a, lwork=int(G["short_lwork"]))
    e = factor_errors(af, tau, G["short_afac"], G["short_tau"])
    print(e)
    assert max(e) <= TOL


def test_batch_problems_against_fixture():
    for i, (m, n) in enumerate(G["batch_shapes"].tolist()):
        af, tau = O.factor(seeded(m, n, i))
        e = factor_errors(af, tau, G[f"batch_{i}_afac"], G[f"batch_{i}_tau"])
        assert max(e) <= TOL, (i, m, n, e)


# -- live reference (built from /root/reference against MKL) -----------------
live = pytest.mark.skipif(not refqr.available(), reason=f"reference QR library unavailable: "
                                                        f"{refqr.why_unavailable()}")


@live
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(300, 200), (129, 300), (700, 257), (1000, 160), (64, 64), (2, 1)])
def test_factor_multq_generateq_live(shape, cplx):
    refqr.set_threads(1)
    m, n = shape
    rng = np.random.default_rng(m * 7 + n)
    a = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
    af_r, tau_r = refqr.factor(a)
    af, tau = O.factor(a)
    e = factor_errors(af, tau, af_r, tau_r)
    assert max(e) <= TOL, e
    k = min(m, n)
    if m >= n:
        q_r = refqr.generateq(af_r, tau_r, n=n, k=max(k - 3, 0))
        q = O.generateq(af_r, tau_r, n=n, k=max(k - 3, 0))
        assert np.max(np.abs(q - q_r)) <= TOL
    for side in "LR":
        for trans in "NC":
            shp = (m, 37) if side == "L" else (37, m)
            c = rng.standard_normal(shp) + (1j * rng.standard_normal(shp) if cplx else 0)
            out_r = refqr.multq(side, trans, af_r, tau_r, c, k=k)
            out = O.multq(side, trans, af_r, tau_r, c, k=k)
            assert np.max(np.abs(out - out_r)) <= TOL * max(1.0, np.max(np.abs(out_r))), (side, trans)


@live
def test_random_matrix_live():
    """MatrixGen::RandomMatrix: the oracle's restatement against the
    reference's, graded singular values (d equal bit for bit; the random
    reflectors applied by BLAS Level 2, equal to rounding), and the seed
    advanced identically."""
    for (m, n, cond, mode) in [(40, 30, 1e6, O.DIAG_GRADED_EXPONENTIALLY), (25, 60, 1e3, O.DIAG_GRADED_LINEARLY),
                               (50, 50, 1e2, O.DIAG_LASTCOND_REMAINING1)]:
        s0 = np.array([3, 17, 29, 41], dtype=np.int32)
        a_r, d_r, s_r = refqr.random_matrix(m, n, cond, mode, 2.0, O.UNIFORM01, s0)
        s1 = s0.copy()
        a = O.random_matrix(m, n, cond, mode, 2.0, O.UNIFORM01, s1)
        assert np.array_equal(s1, s_r)
        assert np.max(np.abs(a - a_r)) <= 1e-14 * max(1.0, np.max(np.abs(a_r)))
        sv = np.linalg.svd(a, compute_uv=False)
        assert np.allclose(np.sort(sv)[::-1], np.sort(np.abs(d_r))[::-1], rtol=1e-12)


def test_factor_panels_equals_factor_leading_panels():
    """factor_panels (the CPU-baseline sample) runs QR::Factor's blocked loop
    for its first P panels only: those columns and taus must equal Factor's
    bit for bit (the operations are the same), and the trailing columns must
    hold the P-panel update."""
    m, n, nb = 600, 400, 32
    a = seeded(m, n, 5)
    af, tau = O.factor(a)
    for P in (1, 3, 7):
        ap = np.asfortranarray(a.copy())
        taup = O.factor_panels(ap, nb, P)
        c = P * nb
        assert np.array_equal(ap[:, :c], af[:, :c]) and np.array_equal(taup[:c], tau[:c]), P
        # the trailing part is Q_P^H A(:, c:), Q_P the product of the first P panels' reflectors
        qp = O.generateq(np.asfortranarray(ap[:, :c]), taup[:c], n=c)
        full = O.multq("L", "C", np.asfortranarray(ap[:, :c]), taup[:c], np.asfortranarray(a[:, c:]))
        assert np.max(np.abs(full - ap[:, c:])) <= 1e-12 * np.max(np.abs(a)), P
        assert qp.shape == (m, c)


# -- LU::Invert (RNP/LA/LU.hpp:146-214) ---------------------------------------
def _lu_case(n, cplx, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    lu, piv, info = O.lu_factor(a)
    assert info == 0
    return a, lu, piv


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 130, 257])
def test_lu_invert_is_the_inverse(n, cplx):
    """The oracle's LU::Invert equals the inverse of A (numpy's, to the
    conditioning) and inv(A) A = I."""
    a, lu, piv = _lu_case(n, cplx, 100 + n)
    inv, st = O.lu_invert(lu, piv)
    assert st == 0
    ref = np.linalg.inv(a)
    kappa = np.linalg.cond(a)
    assert np.max(np.abs(inv - ref)) <= 50 * n * kappa * 2.2e-16 * np.max(np.abs(ref))
    assert np.max(np.abs(inv @ a - np.eye(n))) <= 50 * n * kappa * 2.2e-16


def test_lu_invert_zero_pivot_returns_minus_index():
    """Triangular::Invert's zero test (Triangular.hpp:217-223): the first
    exactly-zero U_ii returns -(i+1) and nothing is modified."""
    rng = np.random.default_rng(4)
    lu = np.triu(rng.standard_normal((6, 6))) + np.tril(rng.standard_normal((6, 6)), -1)
    lu[3, 3] = 0.0
    piv = np.array([2, 1, 5, 3, 4, 5], dtype=np.uint64)
    out, st = O.lu_invert(lu, piv)
    assert st == -4 and np.array_equal(out, lu)


@live
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [5, 64, 65, 130, 200])
def test_lu_invert_live_unblocked_branch(n, cplx):
    """Against the reference's own LU::Invert: its unblocked branch (a
    workspace of n, LU.hpp:170-185) at every n, and its default workspace where
    n <= 64 (the same branch)."""
    refqr.set_threads(1)
    a, lu, piv = _lu_case(n, cplx, n)
    inv, st = O.lu_invert(lu, piv)
    r_inv, r_st = refqr.lu_invert(lu, piv, lwork_cap=n)
    scale = np.max(np.abs(r_inv))
    assert st == r_st == 0
    assert np.max(np.abs(inv - r_inv)) <= 1e-12 * scale
    if n <= 64:
        d_inv, _ = refqr.lu_invert(lu, piv)
        assert np.max(np.abs(inv - d_inv)) <= 1e-12 * scale


@live
def test_reference_blocked_lu_invert_defect():
    """The defect documented for the reference's blocked branch (n > 64 with
    the default workspace: LU.hpp:192-197 saves columns jj, not j + jj, of L):
    its result is not the inverse, while its unblocked branch is."""
    a, lu, piv = _lu_case(100, False, 7)
    d_inv, _ = refqr.lu_invert(lu, piv)
    u_inv, _ = refqr.lu_invert(lu, piv, lwork_cap=100)
    ref = np.linalg.inv(a)
    assert np.max(np.abs(u_inv - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.max(np.abs(d_inv - ref)) > 1e-3 * np.max(np.abs(ref))


# ---------------------------------------------------------------------------
# Inf / NaN in the rows the reference's zero trims skip: the oracle's
# restatement of those trims (Reflector.hpp:36-75, 196-221, 581-623, 742-744)
# against the reference built from its sources.
# ---------------------------------------------------------------------------
import nonfinite_cases as NF  # noqa: E402


@live
@pytest.mark.parametrize("case,seed", NF.cases(), ids=lambda c: c[0] if isinstance(c, tuple) else str(c))
def test_factor_nonfinite_live(case, seed):
    tag, m, n, k0, specials, zc, cplx, lwn = case
    a = NF.build(m, n, k0, specials, zc, cplx, seed)
    refqr.set_threads(1)
    lw = lwn * n if lwn else 0
    af_r, tau_r = refqr.factor(a, lwork_cap=lw) if lw else refqr.factor(a)
    af_o, tau_o = O.factor(a, lwork=lw) if lw else O.factor(a)
    assert not np.isfinite(af_r).all(), "the case must reach a non-finite result"
    NF.compare(af_o, tau_o, af_r, tau_r)
