"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something that does not share its
code:
  * vectors written by the real reference (tests/golden/make_golden.py, from
    RNP's own headers compiled by `make -C oracle ref`), bit for bit;
  * the live reference library oracle/_ref/librnpref.so on fresh random
    inputs, where it has been built;
  * LAPACK through scipy (dgeqrf/zgeqrf, dlarfg/zlarfg, dlarf, dormqr/zunmqr,
    dorgqr/zungqr), which shares RNP's reflector convention except at
    real(alpha) == 0;
  * closed forms (the MCG stream from Python integers, the sign rule, the
    rescale count) and brute force on tiny inputs (products of reflectors).
CPU only; no GPU is touched.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import oracle as O
from metrics import orth_error, qr_residual, r_of

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "rnp_reference_vectors.npz"))
EPS = np.finfo(np.float64).eps


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a.view(np.float64).view(np.uint64)


def assert_bits(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = np.nonzero(bits(a) != bits(b))[0]
    assert bad.size == 0, f"{what}: {bad.size} values differ, first at {bad[:5]}: {a.ravel()[bad[:3]]} vs {b.ravel()[bad[:3]]}"


# ---------------------------------------------------------------------------
# Generators (RNP/Random.hpp:45-320, RNP/MatrixGen.hpp:31-110)
# ---------------------------------------------------------------------------

MCG_A = 33952834046453


def mcg_closed_form(iseed, n):
    """dlaruv-style MCG: state s_{k+1} = a s_k mod 2^48, value 1 - s_{k+1}/2^48."""
    s = (iseed[0] << 36) | (iseed[1] << 24) | (iseed[2] << 12) | iseed[3]
    out = []
    for _ in range(n):
        s = (s * MCG_A) % (1 << 48)
        out.append(1.0 - s / float(1 << 48))
    return np.array(out), [(s >> 36) & 4095, (s >> 24) & 4095, (s >> 12) & 4095, s & 4095]


def test_uniform_closed_form():
    seed = np.array([1, 2, 3, 5], dtype=np.int32)
    x = O.uniform_real_vector(1000, seed)
    ref, seed_after = mcg_closed_form([1, 2, 3, 5], 1000)
    assert_bits(x, ref, "UniformRealVector vs integer MCG")
    assert list(seed) == seed_after


def test_uniform_default_seed_stream():
    O.reset_internal_seed()
    x = O.uniform_real_vector(256)
    assert_bits(x, GOLD["uniform_default_256"], "default-seed stream")
    ref, _ = mcg_closed_form([2398, 691, 782, 721], 256)
    assert_bits(x, ref, "default seed {2398,691,782,721}")


def test_uniform_stream_continues():
    seed = np.array([2398, 691, 782, 721], dtype=np.int32)
    x = np.concatenate([O.uniform_real_vector(100, seed), O.uniform_real_vector(100, seed)])
    assert_bits(x, GOLD["uniform_two_calls"], "two calls continue the stream")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("dist", range(5))
def test_generate_vector(dist, cplx):
    pre = "z" if cplx else "d"
    seed = np.array([1, 2, 3, 5], dtype=np.int32)
    x = O.generate_vector(dist, 300, seed, complex_=cplx)
    ref = GOLD[f"genvec_{pre}_{dist}"]
    assert list(seed) == list(GOLD[f"genvec_{pre}_{dist}_seed_after"])
    if dist in (O.UNIFORM01, O.UNIFORM_11) or (dist == O.UNITDISC and not cplx) or (
            dist == O.UNITCIRCLE and not cplx):
        assert_bits(x, ref, f"GenerateVector dist {dist}")
    else:
        # log1p / sqrt / cos / sin come from the host libm: agree to a few ulp
        np.testing.assert_allclose(x, ref, rtol=4 * EPS, atol=4 * EPS)


def test_generate_diagonal_all_modes():
    for row in GOLD["diag_rows"]:
        mode, reverse, randsign = int(row[0]), int(row[1]), int(row[2])
        seed = np.array([4, 3, 2, 1], dtype=np.int32)
        d = O.generate_diagonal(mode, 17, 1e6, dscale=2.5, dist=O.NORMAL01 if mode == 6 else O.UNIFORM01,
                                reverse=bool(reverse), randsign=bool(randsign), iseed=seed)
        np.testing.assert_allclose(d, row[3:], rtol=4 * EPS, atol=0, err_msg=f"mode {mode}")


def test_graded_diagonal_closed_form():
    d = O.generate_diagonal(O.DIAG_GRADED_EXPONENTIALLY, 2048, 1e6)
    t = np.arange(2048) / 2047.0
    np.testing.assert_allclose(d, 1e6 ** (-t), rtol=64 * EPS)


# ---------------------------------------------------------------------------
# Level 1: Norm2 (BLAS_mix.hpp:725-758), hypot3 (Types.hpp:73-92)
# ---------------------------------------------------------------------------

def test_norm2_against_reference():
    for x, r in zip(GOLD["norm2_in_d"], GOLD["norm2_d"]):
        assert_bits(O.norm2(x), r, "Norm2 real")
    for x, r in zip(GOLD["norm2_in_z"], GOLD["norm2_z"]):
        assert_bits(O.norm2(x), r, "Norm2 complex")
    xs = GOLD["norm2_strided_in"]
    L = O.lib()
    got = L.orc_d_norm2(13, xs.ctypes.data_as(ctypes.c_void_p), 3)
    assert_bits(got, GOLD["norm2_strided_inc3"][0], "Norm2 incx=3")


def test_norm2_scaling_invariant():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1000)
    base = O.norm2(x)
    np.testing.assert_allclose(base, np.linalg.norm(x), rtol=8 * EPS)
    for s in (2.0 ** -1060, 2.0 ** 1000):            # exact power-of-two scalings
        assert O.norm2(x * s) == pytest.approx(base * s, rel=8 * EPS)


def test_hypot3():
    for t, r in zip(GOLD["hypot3_in"], GOLD["hypot3"]):
        assert_bits(O.hypot3(*t), r, f"hypot3{tuple(t)}")
    assert O.hypot3(3.0, 4.0, 12.0) == 13.0


# ---------------------------------------------------------------------------
# Reflector::Generate (Reflector.hpp:117-158)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("pre", ["d", "z"])
def test_generate_against_reference(pre):
    offs = GOLD[f"gen_{pre}_offsets"]
    for i in range(offs.size - 1):
        x = GOLD[f"gen_{pre}_x"][offs[i]:offs[i + 1]]
        al = GOLD[f"gen_{pre}_alpha"][i]
        beta, v, tau = O.generate(complex(al) if pre == "z" else float(al), x)
        assert_bits(beta, GOLD[f"gen_{pre}_beta"][i], f"beta case {i}")
        assert_bits(tau, GOLD[f"gen_{pre}_tau"][i], f"tau case {i}")
        assert_bits(v, GOLD[f"gen_{pre}_v"][offs[i]:offs[i + 1]], f"v case {i}")


def test_generate_sign_rule_differs_from_lapack():
    """real(alpha) == 0: RNP keeps beta positive (Reflector.hpp:131), LAPACK
    dlarfg takes beta = -sign(alpha) |.| with sign(+0) = +, i.e. negative."""
    beta, v, tau = O.generate(0.0, [3.0, 4.0])
    assert beta == 5.0 and tau == 1.0
    np.testing.assert_allclose(v, [-0.6, -0.8], rtol=2 * EPS)   # v = x * (1/(alpha - beta))
    lb, lx, ltau = lapack.dlarfg(3, 0.0, np.array([3.0, 4.0]))
    assert lb == -5.0 and ltau == 1.0


def test_generate_matches_lapack_dlarfg():
    rng = np.random.default_rng(11)
    for n in (2, 7, 33, 200):
        for _ in range(5):
            x = rng.standard_normal(n - 1)
            al = rng.standard_normal()
            beta, v, tau = O.generate(al, x)
            lb, lx, ltau = lapack.dlarfg(n, al, x.copy())
            assert beta == pytest.approx(lb, rel=4 * EPS)
            assert tau == pytest.approx(ltau, rel=8 * EPS)
            np.testing.assert_allclose(v, lx, rtol=16 * EPS, atol=16 * EPS * np.abs(lx).max())


def test_generate_matches_lapack_zlarfg():
    rng = np.random.default_rng(12)
    for n in (2, 9, 64):
        for _ in range(5):
            x = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
            al = complex(rng.standard_normal(), rng.standard_normal())
            beta, v, tau = O.generate(al, x)
            lb, lx, ltau = lapack.zlarfg(n, al, x.copy())
            assert abs(beta - lb) <= 8 * EPS * abs(lb)
            assert abs(tau - ltau) <= 16 * EPS * abs(ltau)
            np.testing.assert_allclose(v, lx, rtol=32 * EPS, atol=32 * EPS * np.abs(lx).max())


@pytest.mark.parametrize("cplx", [False, True])
def test_generate_invariants(cplx):
    """H^H [alpha; x] = [beta; 0], |tau - 1| <= 1 (Reflector.hpp:97), H unitary."""
    rng = np.random.default_rng(5)
    for n in (2, 5, 40):
        x = rng.standard_normal(n - 1) + (1j * rng.standard_normal(n - 1) if cplx else 0)
        al = complex(rng.standard_normal(), rng.standard_normal()) if cplx else rng.standard_normal()
        beta, v, tau = O.generate(al, x)
        w = np.concatenate([[1.0], v])
        H = np.eye(n) - tau * np.outer(w, w.conj())
        y = H.conj().T @ np.concatenate([[al], x])
        np.testing.assert_allclose(y, np.concatenate([[beta], np.zeros(n - 1)]), atol=32 * EPS * abs(beta))
        assert abs(np.imag(beta)) == 0.0
        assert abs(tau - 1) <= 1 + 4 * EPS
        np.testing.assert_allclose(H.conj().T @ H, np.eye(n), atol=32 * EPS)


def test_generate_rescale_path():
    """|beta| < 2^-971 takes the rescale loop; the result equals the
    unscaled closed form beta = -sign(alpha) hypot(alpha, |x|)."""
    al, x = 1e-300, np.array([1e-300, 3e-300])
    beta, v, tau = O.generate(al, x)
    exact = -np.sqrt(11.0) * 1e-300
    assert beta == pytest.approx(exact, rel=4 * EPS)
    assert tau == pytest.approx((exact - al) / exact, rel=8 * EPS)
    np.testing.assert_allclose(v, x / (al - exact), rtol=8 * EPS)


def test_generate_tau_zero_cases():
    beta, v, tau = O.generate(2.0, [0.0, 0.0])
    assert tau == 0.0 and beta == 2.0
    beta, v, tau = O.generate(2j, np.zeros(2, dtype=complex))   # imaginary alpha: not identity
    assert tau != 0 and np.imag(beta) == 0.0 and beta == 2.0


def _ref_lib():
    path = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "librnpref.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    for pre in ("d", "z"):
        getattr(L, f"ref_{pre}_generate").argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_size_t, ctypes.c_void_p]
    return L


@pytest.mark.parametrize("pre", ["d", "z"])
def test_generate_live_reference(pre):
    """Random columns of every magnitude band through the compiled reference."""
    L = _ref_lib()
    if L is None:
        pytest.skip("oracle/_ref/librnpref.so not built (reference sources absent)")
    rng = np.random.default_rng(99)
    dt = np.complex128 if pre == "z" else np.float64
    f = getattr(L, f"ref_{pre}_generate")
    for trial in range(400):
        n = int(rng.integers(1, 80))
        scale = 10.0 ** rng.choice([-320, -305, -200, -160, 0, 150, 300])
        x = rng.standard_normal(n - 1) * scale
        al = rng.standard_normal() * scale
        if pre == "z":
            x = x + 1j * rng.standard_normal(n - 1) * scale
            al = complex(al, rng.standard_normal() * scale)
        if trial % 17 == 0:
            al = 0.0 if pre == "d" else complex(0.0, al.imag)
        if trial % 23 == 0:
            x[:] = 0
        xa = np.array(x, dtype=dt, copy=True)
        ala = np.array([al], dtype=dt)
        t = np.zeros(1, dtype=dt)
        f(n, ala.ctypes.data_as(ctypes.c_void_p), xa.ctypes.data_as(ctypes.c_void_p), 1,
          t.ctypes.data_as(ctypes.c_void_p))
        beta, v, tau = O.generate(complex(al) if pre == "z" else float(al), x)
        assert_bits(beta, ala[0], f"beta trial {trial}")
        assert_bits(tau, t[0], f"tau trial {trial}")
        assert_bits(v, xa, f"v trial {trial}")


# ---------------------------------------------------------------------------
# Reflector::Apply (191-360), GenerateBlockTr (573-677), ApplyBlock (725-815)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side", ["L", "R"])
def test_apply_matches_lapack_dlarf(cplx, side):
    rng = np.random.default_rng(21)
    m, n = 13, 9
    c = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
    lv = m if side == "L" else n
    v = rng.standard_normal(lv) + (1j * rng.standard_normal(lv) if cplx else 0)
    v[0] = 1.0
    tau = 1.3 + (0.4j if cplx else 0)
    got = O.apply(side, 1, v, tau, c)
    H = np.eye(lv) - tau * np.outer(v, v.conj())
    want = H @ c if side == "L" else c @ H
    np.testing.assert_allclose(got, want, atol=64 * EPS * np.abs(want).max())
    f = lapack.zlarf if cplx else lapack.dlarf
    lw = f(v, tau, np.asfortranarray(c), np.zeros(max(m, n), dtype=c.dtype), side=side)
    np.testing.assert_allclose(got, lw, atol=64 * EPS * np.abs(want).max())


def _reflector_product(v, tau):
    n, k = v.shape
    Q = np.eye(n, dtype=v.dtype)
    for i in range(k):
        Q = Q @ (np.eye(n) - tau[i] * np.outer(v[:, i], v[:, i].conj()))
    return Q


def _unit_lower(rng, n, k, cplx):
    v = rng.standard_normal((n, k)) + (1j * rng.standard_normal((n, k)) if cplx else 0)
    v = np.tril(v, -1)
    v[np.arange(k), np.arange(k)] = 1.0
    return np.asfortranarray(v)


@pytest.mark.parametrize("cplx", [False, True])
def test_block_tr_brute_force(cplx):
    """H_0 H_1 ... H_{k-1} = I - V T V^H, T upper with diag tau; a tau = 0
    column leaves its T column zero (Reflector.hpp:584-587)."""
    rng = np.random.default_rng(31)
    for n, k in ((6, 3), (20, 8), (9, 9), (40, 32)):
        v = _unit_lower(rng, n, k, cplx)
        tau = 1.0 + rng.uniform(0, 1, k) * (np.exp(1j * rng.uniform(0, 6, k)) if cplx else 1)
        if k > 2:
            tau[1] = 0.0
        t = O.generate_block_tr(v, tau)
        assert np.all(np.tril(t, -1) == 0)
        np.testing.assert_array_equal(np.diag(t), tau)
        if k > 2:
            assert np.all(t[:, 1] == 0)
        Q = _reflector_product(v, tau)
        np.testing.assert_allclose(np.eye(n) - v @ t @ v.conj().T, Q, atol=64 * k * EPS)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_apply_block_equals_sequential(cplx, side, trans):
    rng = np.random.default_rng(41)
    m, n, k = 37, 29, 8
    nq = m if side == "L" else n
    v = _unit_lower(rng, nq, k, cplx)
    tau = np.array([O.generate(complex(x) if cplx else x, rng.standard_normal(3))[2]
                    for x in rng.standard_normal(k)])
    t = O.generate_block_tr(v, tau)
    c = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
    got = O.apply_block(side, trans, v, t, c)
    Q = _reflector_product(v, tau)
    op = Q if trans == "N" else Q.conj().T
    want = op @ c if side == "L" else c @ op
    np.testing.assert_allclose(got, want, atol=128 * k * EPS * np.abs(c).max())


def test_last_nonzero_limits():
    a = np.zeros((7, 5))
    assert O.last_nonzero_column_limit(a) == 0 and O.last_nonzero_row_limit(a) == 0
    a[3, 2] = 1.0
    assert O.last_nonzero_column_limit(a) == 3
    assert O.last_nonzero_row_limit(a) == 4
    a[6, 4] = 2.0
    assert O.last_nonzero_column_limit(a) == 5 and O.last_nonzero_row_limit(a) == 7


# ---------------------------------------------------------------------------
# QR::Factor / MultQ / GenerateQ (QR.hpp:78-570)
# ---------------------------------------------------------------------------

def _rand(rng, m, n, cplx):
    a = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
    return np.asfortranarray(a)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n", [(170, 160), (166, 170), (512, 256), (300, 300), (129, 5)])
def test_factor_matches_lapack_geqrf(m, n, cplx):
    rng = np.random.default_rng(m * 7 + n)
    a = _rand(rng, m, n, cplx)
    af, tau = O.factor(a)
    geqrf = lapack.zgeqrf if cplx else lapack.dgeqrf
    lq, ltau, _, info = geqrf(a.copy(order="F"))
    assert info == 0
    k = min(m, n)
    np.testing.assert_allclose(tau, ltau[:k], rtol=0, atol=1e-12)
    R, LR = np.triu(af[:k]), np.triu(lq[:k])
    scale = np.abs(LR).max(axis=0)
    assert np.max(np.abs(R - LR) / scale) < 1e-12
    V, LV = np.tril(af, -1), np.tril(lq, -1)
    assert np.max(np.abs(V - LV)) < 1e-10


@pytest.mark.parametrize("cplx", [False, True])
def test_factor_unblocked_equals_blocked(cplx):
    rng = np.random.default_rng(7)
    a = _rand(rng, 300, 200, cplx)
    af1, t1 = O.factor(a)
    af2, t2 = O.factor(a, unblocked=True)
    np.testing.assert_allclose(af1, af2, atol=1e-12 * np.abs(af2).max())
    np.testing.assert_allclose(t1, t2, atol=1e-13)


def test_factor_workspace_query_and_degrade():
    assert O.factor_query(512, 256) == 32 * 256
    assert O.factor_query(8192, 8192) == 32 * 8192
    rng = np.random.default_rng(8)
    a = _rand(rng, 400, 300, False)
    af, tau = O.factor(a)
    for lw in (16 * 300, 2 * 300, 400):   # nb degrades to 16, 2, then unblocked
        af2, tau2 = O.factor(a, lwork=lw)
        np.testing.assert_allclose(af2, af, atol=1e-12 * np.abs(af).max())
        np.testing.assert_allclose(tau2, tau, atol=1e-13)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n", [(170, 160), (166, 170)])
def test_factor_reference_test_identities(m, n, cplx):
    """The checks of tests/test_qr_tall.cpp:26-227 and test_qr_fat.cpp, in
    units of m n eps (the reference prints 0.1-3 for these)."""
    rng = np.random.default_rng(1)
    a = _rand(rng, m, n, cplx)
    af, tau = O.factor(a)
    k = min(m, n)
    Q = O.generateq(af[:, :k], tau, n=k)
    R = r_of(af)[:k]
    unit = m * n * EPS
    assert np.linalg.norm(Q @ R - a) / np.linalg.norm(a) / unit < 30
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(k)) / unit < 30
    for side, trans in (("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")):
        c = _rand(rng, m if side == "L" else 7, 7 if side == "L" else m, cplx)
        got = O.multq(side, trans, af[:, :k], tau, c)
        Qf = O.generateq(np.hstack([af[:, :k], np.zeros((m, m - k), dtype=af.dtype)]), tau, n=m, k=k)
        op = Qf if trans == "N" else Qf.conj().T
        want = op @ c if side == "L" else c @ op
        assert np.linalg.norm(got - want) / np.linalg.norm(c) / unit < 30


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_multq_matches_lapack(cplx, side, trans):
    rng = np.random.default_rng(17)
    m, k = 200, 150
    a = _rand(rng, m, k, cplx)
    af, tau = O.factor(a)
    c = _rand(rng, m if side == "L" else 45, 45 if side == "L" else m, cplx)
    got = O.multq(side, trans, af, tau, c)
    f = lapack.zunmqr if cplx else lapack.dormqr
    ltrans = trans if cplx else ("T" if trans == "C" else "N")
    lw = f(side, ltrans, af, tau, c.copy(order="F"), -1)[1]
    want, _, info = f(side, ltrans, af, tau, c.copy(order="F"), int(np.real(lw[0])))
    assert info == 0
    np.testing.assert_allclose(got, want, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("cplx", [False, True])
def test_generateq_matches_lapack(cplx):
    rng = np.random.default_rng(19)
    m, n = 300, 200
    af, tau = O.factor(_rand(rng, m, n, cplx))
    got = O.generateq(af, tau)
    f = lapack.zungqr if cplx else lapack.dorgqr
    want, _, info = f(af.copy(order="F"), tau)
    assert info == 0
    np.testing.assert_allclose(got, want, atol=1e-12)
    for k in (0, 1, 37):                      # fewer reflectors than columns
        q = O.generateq(af, tau[:k], k=k)
        w, _, _ = f(af.copy(order="F"), np.concatenate([tau[:k], np.zeros(n - k, dtype=tau.dtype)]))
        np.testing.assert_allclose(q, w, atol=1e-12)


def test_multq_and_generateq_queries():
    L = O.lib()
    lw = ctypes.c_size_t(0)
    L.orc_d_multq(b"L", b"C", 100, 40, 30, None, 100, None, None, 100, ctypes.byref(lw), None, None)
    assert lw.value == 32 * 40 + 32 * 32
    lw = ctypes.c_size_t(0)
    L.orc_d_generateq(100, 40, 30, None, 100, None, ctypes.byref(lw), None, None)
    assert lw.value == 40 * 32


@pytest.mark.parametrize("cplx", [False, True])
def test_factor_edge_inputs(cplx):
    dt = np.complex128 if cplx else np.float64
    for a in (np.zeros((9, 6), dtype=dt), np.eye(8, 5, dtype=dt), np.ones((1, 4), dtype=dt),
              np.ones((5, 1), dtype=dt) * 1e-310):
        af, tau = O.factor(np.asfortranarray(a))
        assert np.all(np.isfinite(af)) and np.all(np.isfinite(tau))
        if not np.any(a):
            assert np.all(tau == 0) and np.all(af == 0)
    rng = np.random.default_rng(2)
    a = _rand(rng, 60, 40, cplx)
    a[:, 10] = 0.0
    a[:, 20] = a[:, 3]                        # rank deficient
    af, tau = O.factor(a)
    k = 40
    Q = O.generateq(af, tau)
    assert qr_residual(a, Q, r_of(af)[:k]) <= 10
    assert orth_error(Q) <= 10


def test_ls_solve_matches_lstsq():
    rng = np.random.default_rng(4)
    a = _rand(rng, 120, 30, False)
    b = rng.standard_normal((120, 3))
    x = O.ls_solve(a, b)
    np.testing.assert_allclose(x, np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("N", "C"), ("T", "N")])
def test_gemm(ta, tb):
    rng = np.random.default_rng(6)
    a = rng.standard_normal((70, 50)) + 1j * rng.standard_normal((70, 50))
    b = rng.standard_normal((50, 33)) + 1j * rng.standard_normal((50, 33))
    A = a if ta == "N" else (a.conj().T.copy() if ta == "C" else a.T.copy())
    B = b if tb == "N" else (b.conj().T.copy() if tb == "C" else b.T.copy())
    got = O.gemm(ta, tb, 2.0 - 1j, A, B)
    np.testing.assert_allclose(got, (2.0 - 1j) * (a @ b), atol=1e-12 * np.abs(a @ b).max())


# ---------------------------------------------------------------------------
# LQ (RNP/LA/LQ.hpp).  LAPACK's gelq2 is conj(row) + larfg + larf from the
# right, i.e. QR of A^H: the oracle's row-wise restatement must reproduce
# LAPACK geqrf on A^H (R^H = L, conj of V in the rows, same tau).
# ---------------------------------------------------------------------------

LQ_SHAPES = [(5, 9), (9, 5), (40, 70), (70, 40), (33, 33), (1, 1), (1, 6), (6, 1), (130, 200)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", LQ_SHAPES)
def test_lq_factor_matches_lapack_qr_of_adjoint(shape, cplx):
    m, n = shape
    rng = np.random.default_rng(m * 1000 + n)
    a = _rand(rng, m, n, cplx)
    af, tau = O.lq_factor(a)
    k = min(m, n)
    qr, tl, _, info = (lapack.zgeqrf if cplx else lapack.dgeqrf)(a.conj().T.copy())
    assert info == 0
    L = np.tril(af)[:, :k]
    s = max(np.abs(a).max(), 1e-300)
    assert np.abs(L - np.triu(qr)[:k].conj().T).max() <= 1e-13 * s * max(m, n)
    assert np.abs(np.triu(af, 1)[:k] - np.tril(qr, -1)[:, :k].conj().T).max() <= 1e-12
    np.testing.assert_allclose(tau, tl[:k], rtol=0, atol=1e-13)


@pytest.mark.parametrize("cplx", [False, True])
def test_lq_brute_force_reflectors(cplx):
    """Q from GenerateQ equals (H_0 H_1 ... H_{k-1})^H built from the stored
    rows (v_i = conj of row i, unit at i) and tau; A = L Q, Q Q^H = I."""
    rng = np.random.default_rng(8)
    m, n = 6, 9
    a = _rand(rng, m, n, cplx)
    af, tau = O.lq_factor(a)
    H = np.eye(n, dtype=af.dtype)
    for i in range(m):
        v = np.zeros(n, dtype=af.dtype)
        v[i] = 1.0
        v[i + 1:] = np.conj(af[i, i + 1:])
        H = H @ (np.eye(n) - tau[i] * np.outer(v, v.conj()))
    Q = O.lq_generateq(af, tau)
    np.testing.assert_allclose(Q, H.conj().T[:m], atol=1e-14)
    np.testing.assert_allclose(np.tril(af)[:, :m] @ Q, a, atol=1e-13)
    np.testing.assert_allclose(Q @ Q.conj().T, np.eye(m), atol=1e-14)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_lq_multq_against_explicit_q(side, trans, cplx):
    """MultQ (LQ.hpp:195-196: Q C, Q^H C, C Q, C Q^H) against the explicit Q
    of GenerateQ.  k <= 32 takes the reference's unblocked loop; its side "L"
    branch of Reflector::Apply (Reflector.hpp:258-263) updates column 0 only,
    so the restatement follows the documented product instead."""
    rng = np.random.default_rng(12)
    for k, nq in ((5, 11), (20, 40)):
        a = _rand(rng, k, nq, cplx)
        af, tau = O.lq_factor(a)
        full = np.zeros((nq, nq), dtype=af.dtype, order="F")
        full[:k] = af
        Q = O.lq_generateq(full, tau, m=nq, k=k)
        c = _rand(rng, nq, 7, cplx) if side == "L" else _rand(rng, 7, nq, cplx)
        got = O.lq_multq(side, trans, af, tau, c, k=k)
        op = Q if trans == "N" else Q.conj().T
        want = op @ c if side == "L" else c @ op
        np.testing.assert_allclose(got, want, atol=1e-13 * max(1.0, np.abs(want).max()))


@pytest.mark.parametrize("cplx", [False, True])
def test_lq_sign_rule_and_edges(cplx):
    """Re(alpha) == 0 gives beta > 0 (Reflector.hpp:131); a zero row takes
    tau = 0; tiny and huge rows take the scaled paths."""
    rng = np.random.default_rng(14)
    a = _rand(rng, 8, 12, cplx)
    a[0, 0] = 0.0
    a[2, :] = 0.0
    a[3, :] *= 1e-300
    a[4, :] *= 1e300
    af, tau = O.lq_factor(a)
    assert af[0, 0].real > 0
    assert tau[2] == 0
    # LAPACK's sign at alpha == 0 differs, and with the zero row that changes
    # the later pivots; the oracle's own column-wise QR of A^H (RNP's rule,
    # pinned above) is the comparison: same factors, conjugate-transposed.
    qf, qt = O.factor(a.conj().T)
    k = 8
    np.testing.assert_allclose(np.diag(af).real, np.diag(qf).real, rtol=1e-13)
    np.testing.assert_allclose(tau, qt, rtol=0, atol=1e-13)
    sc = np.abs(np.tril(af)[:, :k]).max(axis=1, keepdims=True)
    sc[sc == 0] = 1.0
    assert np.max(np.abs(np.tril(af)[:, :k] - np.triu(qf)[:k].conj().T) / sc) <= 1e-13


def test_lq_workspace_queries():
    assert O.lq_factor_query(50, 70) == 32 * 70
    assert O.lq_factor_query(0, 70) == 0


# ---------------------------------------------------------------------------
# QL (QL.hpp) and RQ (RQ.hpp).  Pins: LAPACK's QR of the matrix with rows
# and columns reversed (QL: the reversal of A = Q L is a QR factorization
# with the same reflectors reversed, so LAPACK's independent geqrf pins every
# stored entry), LAPACK gerqf/orgrq directly for RQ, the explicit reflector
# products (brute force) for MultQ and GenerateQ, and A = Q L / A = R Q.
# ---------------------------------------------------------------------------

QLRQ_SHAPES = [(5, 9), (9, 5), (40, 70), (70, 40), (33, 33), (1, 1), (1, 6), (6, 1), (130, 200)]


def _rot(x):
    return np.asfortranarray(x[::-1, ::-1])


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", QLRQ_SHAPES)
def test_ql_factor_matches_lapack_qr_of_reversed(shape, cplx):
    m, n = shape
    rng = np.random.default_rng(m * 977 + n)
    a = _rand(rng, m, n, cplx)
    af, tau = O.ql_factor(a)
    k = min(m, n)
    qr, tq, _, info = (lapack.zgeqrf if cplx else lapack.dgeqrf)(_rot(a))
    assert info == 0
    want = _rot(qr)
    s = max(np.abs(a).max(), 1e-300)
    # L: rows m-k.., on and below the diagonal that ends at (m-1, n-1)
    lmask = np.zeros((m, n), bool)
    for i in range(m):
        for j in range(n):
            lmask[i, j] = (j - (n - k)) <= (i - (m - k)) and i >= m - k or (j < n - k)
    vmask = ~lmask
    assert np.abs((af - want)[lmask]).max() <= 1e-13 * s * max(m, n)
    if vmask.any():
        assert np.abs((af - want)[vmask]).max() <= 1e-12
    np.testing.assert_allclose(tau, tq[:k][::-1], rtol=0, atol=1e-13)


def _ql_explicit_q(af, tau, k):
    """H[k-1] ... H[0] (QL.hpp:59-64; the order MultQ_unblocked applies),
    H[i] = I - tau_i v v^H, v stored in column n-k+i above row nq-k+i, 1 there."""
    nq, n = af.shape
    Q = np.eye(nq, dtype=af.dtype)
    for i in range(k):
        v = np.zeros(nq, dtype=af.dtype)
        r = nq - k + i
        v[:r] = af[:r, n - k + i]
        v[r] = 1.0
        Q = (np.eye(nq) - tau[i] * np.outer(v, v.conj())) @ Q
    return Q


@pytest.mark.parametrize("cplx", [False, True])
def test_ql_brute_force_reflectors(cplx):
    """GenerateQ equals the last n columns of H[k-1] ... H[0] from the stored
    vectors; A = Q L with L the bottom n x n lower triangle; Q^H Q = I."""
    rng = np.random.default_rng(21)
    m, n = 11, 6
    a = _rand(rng, m, n, cplx)
    af, tau = O.ql_factor(a)
    H = _ql_explicit_q(af, tau, n)
    Q = O.ql_generateq(af.copy(), tau)
    np.testing.assert_allclose(Q, H[:, m - n:], atol=1e-14)
    L = np.tril(af[m - n:, :])
    np.testing.assert_allclose(Q @ L, a, atol=1e-13)
    np.testing.assert_allclose(Q.conj().T @ Q, np.eye(n), atol=1e-14)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_ql_multq_against_explicit_q(side, trans, cplx):
    rng = np.random.default_rng(22)
    for nq, k in ((11, 5), (40, 20)):
        a = _rand(rng, nq, k, cplx)
        af, tau = O.ql_factor(a)
        Q = _ql_explicit_q(af, tau, k)
        c = _rand(rng, nq, 7, cplx) if side == "L" else _rand(rng, 7, nq, cplx)
        got = O.ql_multq(side, trans, af, tau, c, k=k)
        op = Q if trans == "N" else Q.conj().T
        want = op @ c if side == "L" else c @ op
        np.testing.assert_allclose(got, want, atol=1e-13 * max(1.0, np.abs(want).max()))


@pytest.mark.parametrize("cplx", [False, True])
def test_ql_generateq_partial_k(cplx):
    """k < n: the first n-k columns start as unit columns (QL.hpp:455-460)."""
    rng = np.random.default_rng(23)
    m, n, k = 12, 7, 4
    a = _rand(rng, m, k, cplx)
    af, tau = O.ql_factor(a)
    full = np.zeros((m, n), dtype=af.dtype, order="F")
    full[:, n - k:] = af
    Q = O.ql_generateq(full, tau, k=k)
    H = _ql_explicit_q(af, tau, k)
    np.testing.assert_allclose(Q, H[:, m - n:], atol=1e-14)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", QLRQ_SHAPES)
def test_rq_factor_matches_lapack_gerqf(shape, cplx):
    m, n = shape
    rng = np.random.default_rng(m * 991 + n)
    a = _rand(rng, m, n, cplx)
    af, tau = O.rq_factor(a)
    k = min(m, n)
    rq, tl, _, info = (lapack.zgerqf if cplx else lapack.dgerqf)(a.copy())
    assert info == 0
    s = max(np.abs(a).max(), 1e-300)
    rmask = np.zeros((m, n), bool)
    for i in range(m):
        for j in range(n):
            rmask[i, j] = (i < m - k) or (j - (n - k) >= i - (m - k))
    assert np.abs((af - rq)[rmask]).max() <= 1e-13 * s * max(m, n)
    if (~rmask).any():
        assert np.abs((af - rq)[~rmask]).max() <= 1e-12
    np.testing.assert_allclose(tau, tl[:k], rtol=0, atol=1e-13)


@pytest.mark.parametrize("cplx", [False, True])
def test_rq_generateq_matches_lapack_orgrq(cplx):
    """GenerateQ against LAPACK orgrq/ungrq for k = m and for the last k < m
    reflector rows (the leading rows start as unit rows, RQ.hpp:465-474)."""
    rng = np.random.default_rng(24)
    for m, n, k in ((6, 11, 6), (6, 11, 4), (30, 50, 30), (30, 50, 17)):
        a = _rand(rng, m, n, cplx)
        af, tau = O.rq_factor(a)
        inp = np.zeros((m, n), dtype=af.dtype, order="F")
        inp[m - k:] = af[m - k:]
        got = O.rq_generateq(inp.copy(), tau[m - k:], k=k)
        qw, _, info = (lapack.zungrq if cplx else lapack.dorgrq)(inp.copy(), tau[m - k:])
        assert info == 0
        np.testing.assert_allclose(got, qw, atol=1e-13)
        np.testing.assert_allclose(got @ got.conj().T, np.eye(m), atol=1e-13)
        if k == m:
            np.testing.assert_allclose(np.triu(af[:, n - m:]) @ got, a, atol=1e-12)


def _rq_explicit_q(af, tau, k):
    """Q = H[0]^H ... H[k-1]^H with v_i = conj of the stored row m-k+i left of
    column nq-k+i, 1 there (RQ.hpp:57-66; LAPACK's dgerqf convention)."""
    m, nq = af.shape
    Q = np.eye(nq, dtype=af.dtype)
    for i in range(k):
        v = np.zeros(nq, dtype=af.dtype)
        c = nq - k + i
        v[:c] = np.conj(af[m - k + i, :c])
        v[c] = 1.0
        Q = Q @ (np.eye(nq) - tau[i] * np.outer(v, v.conj())).conj().T
    return Q


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_rq_multq_against_explicit_q(side, trans, cplx):
    rng = np.random.default_rng(25)
    for k, nq in ((5, 11), (20, 40)):
        a = _rand(rng, k, nq, cplx)
        af, tau = O.rq_factor(a)
        Q = _rq_explicit_q(af, tau, k)
        np.testing.assert_allclose(np.triu(af[:, nq - k:]) @ Q[nq - k:], a, atol=1e-12)
        c = _rand(rng, nq, 7, cplx) if side == "L" else _rand(rng, 7, nq, cplx)
        got = O.rq_multq(side, trans, af, tau, c, k=k)
        op = Q if trans == "N" else Q.conj().T
        want = op @ c if side == "L" else c @ op
        np.testing.assert_allclose(got, want, atol=1e-13 * max(1.0, np.abs(want).max()))


@pytest.mark.parametrize("cplx", [False, True])
def test_ql_rq_sign_rule_and_edges(cplx):
    """Re(alpha) == 0 gives beta > 0 (Reflector.hpp:131); a zero column (row)
    takes tau = 0; tiny and huge columns (rows) take the scaled paths."""
    rng = np.random.default_rng(26)
    a = _rand(rng, 12, 8, cplx)
    a[11, 7] = 0.0          # QL's first alpha
    a[:, 5] = 0.0           # a zero column: its reflector is H = I
    a[:, 3] *= 1e-300
    a[:, 2] *= 1e300
    af, tau = O.ql_factor(a)
    assert af[11, 7].real > 0
    assert tau[5] == 0
    # the oracle's own QR of the reversed matrix (RNP's sign rule) agrees
    qf, qt = O.factor(_rot(a))
    np.testing.assert_allclose(tau, qt[::-1], rtol=0, atol=1e-13)
    L = np.tril(af[4:, :])
    Lw = np.tril(_rot(qf)[4:, :])
    sc = np.abs(Lw).max(axis=0, keepdims=True)
    sc[sc == 0] = 1.0
    assert np.max(np.abs(L - Lw) / sc) <= 1e-13
    b = _rand(rng, 8, 12, cplx)
    b[7, 11] = 0.0
    b[5, :] = 0.0
    af, tau = O.rq_factor(b)
    assert af[7, 11].real > 0
    assert tau[5] == 0


def test_ql_rq_workspace_queries():
    assert O.ql_query("factor", 50, 70) == 32 * 70
    assert O.rq_query("factor", 50, 70) == 32 * 50
    assert O.ql_query("factor", 0, 70) == 0
    assert O.ql_query("multq", "L", "N", 40, 9, 5) == 32 * 9 + 32 * 32
    assert O.rq_query("multq", "R", "C", 40, 9, 5) == 32 * 40 + 32 * 32
    assert O.ql_query("generateq", 40, 30, 20) == 30 * 32
    assert O.rq_query("generateq", 30, 40, 20) == 30 * 32


# ---------------------------------------------------------------------------
# Hessenberg::MultQ / GenerateQ (Hessenberg.hpp:478-588).  Inputs are LAPACK
# gehrd reductions (RNP's 0-based ilo and exclusive ihi are LAPACK's ilo-1
# and ihi); pins: LAPACK orghr/unghr for GenerateQ, Q^H A Q upper Hessenberg,
# and the four MultQ products against the explicit Q.
# ---------------------------------------------------------------------------

HESS_CASES = [(12, 0, 12), (12, 2, 9), (40, 0, 40), (40, 5, 31), (7, 3, 4), (1, 0, 1)]


def _gehrd(rng, n, ilo, ihi, cplx):
    a = _rand(rng, n, n, cplx)
    # outside [ilo, ihi) the matrix must already be triangular for gehrd's
    # contract; zero those entries below the diagonal
    for j in range(n):
        for i in range(j + 1, n):
            if j < ilo or i >= ihi:
                a[i, j] = 0.0
    ht, tau, info = (lapack.zgehrd if cplx else lapack.dgehrd)(a, lo=ilo, hi=max(ihi - 1, 0))
    assert info == 0
    return a, np.asfortranarray(ht), tau


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n,ilo,ihi", HESS_CASES)
def test_hess_generateq_matches_lapack_orghr(n, ilo, ihi, cplx):
    rng = np.random.default_rng(n * 31 + ilo)
    a, ht, tau = _gehrd(rng, n, ilo, ihi, cplx)
    got = O.hess_generateq(ht.copy(), tau, ilo, ihi)
    if n == 1:
        qw = np.ones((1, 1))   # scipy's wrapper rejects the empty tau; Q = I
    else:
        qw, info = (lapack.zunghr if cplx else lapack.dorghr)(ht.copy(), tau, lo=ilo, hi=max(ihi - 1, 0))
        assert info == 0
    np.testing.assert_allclose(got, qw, atol=1e-13)
    h = got.conj().T @ a @ got
    assert np.abs(np.tril(h, -2)).max(initial=0.0) <= 1e-12 * max(1.0, np.abs(a).max())
    np.testing.assert_allclose(got.conj().T @ got, np.eye(n), atol=1e-13)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_hess_multq_against_explicit_q(side, trans, cplx):
    rng = np.random.default_rng(41)
    for n, ilo, ihi in ((12, 2, 9), (40, 0, 40)):
        a, ht, tau = _gehrd(rng, n, ilo, ihi, cplx)
        Q = O.hess_generateq(ht.copy(), tau, ilo, ihi)
        c = _rand(rng, n, 7, cplx) if side == "L" else _rand(rng, 7, n, cplx)
        got = O.hess_multq(side, trans, ht, tau, c, ilo, ihi)
        op = Q if trans == "N" else Q.conj().T
        want = op @ c if side == "L" else c @ op
        np.testing.assert_allclose(got, want, atol=1e-13 * max(1.0, np.abs(want).max()))


def test_hess_workspace_queries():
    assert O.hess_query("generateq", 10, 2, 8) == 5 * 32
    assert O.hess_query("multq", "L", "N", 10, 5, 2, 8) == 32 * 5 + 32 * 32
    assert O.hess_query("multq", "R", "C", 6, 10, 2, 8) == 32 * 6 + 32 * 32


# ---------------------------------------------------------------------------
# LU (LU.hpp): LAPACK getrf/getrs (same partial pivoting rule, so the same
# pivots), P A = L U, and the reference's return value arithmetic.
# ---------------------------------------------------------------------------

LU_SHAPES = [(5, 5), (7, 5), (5, 7), (64, 64), (65, 65), (200, 200), (300, 150), (150, 300), (1, 1), (1, 9), (9, 1)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", LU_SHAPES)
def test_lu_factor_matches_lapack_getrf(shape, cplx):
    m, n = shape
    rng = np.random.default_rng(m * 7 + n)
    a = _rand(rng, m, n, cplx)
    lu, piv, info = O.lu_factor(a)
    lw, pw, iw = (lapack.zgetrf if cplx else lapack.dgetrf)(a.copy())
    assert info == 0 and iw == 0
    k = min(m, n)
    assert np.array_equal(piv, pw[:k].astype(np.uint64))
    assert np.abs(lu - lw).max() <= 1e-12 * max(1.0, np.abs(a).max()) * max(m, n)
    # P A = L U
    L = np.tril(lu[:, :k], -1) + np.eye(m, k)
    U = np.triu(lu[:k, :])
    pa = a.copy()
    for i in range(k):
        pa[[i, int(piv[i])]] = pa[[int(piv[i]), i]]
    np.testing.assert_allclose(L @ U, pa, atol=1e-12 * max(1.0, np.abs(a).max()) * max(m, n))


def test_lu_unblocked_equals_blocked_to_rounding():
    rng = np.random.default_rng(3)
    a = _rand(rng, 150, 150, False)
    lb, pb, _ = O.lu_factor(a)
    lu_, pu, _ = O.lu_factor(a, unblocked=True)
    assert np.array_equal(pb, pu)
    assert np.abs(lb - lu_).max() <= 1e-12


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_lu_solve_matches_lapack_getrs(trans, cplx):
    rng = np.random.default_rng(5)
    for n in (6, 130):
        a = _rand(rng, n, n, cplx)
        b = _rand(rng, n, 4, cplx)
        lu, piv, _ = O.lu_factor(a)
        x = O.lu_solve(trans, lu, piv, b)
        op = a if trans == "N" else (a.T if trans == "T" else a.conj().T)
        np.testing.assert_allclose(op @ x, b, atol=1e-11 * np.abs(b).max())
        lw, pw, _ = (lapack.zgetrf if cplx else lapack.dgetrf)(a.copy())
        tcode = {"N": 0, "T": 1, "C": 2}[trans]
        xw, info = (lapack.zgetrs if cplx else lapack.dgetrs)(lw, pw, b.copy(), trans=tcode)
        assert info == 0
        np.testing.assert_allclose(x, xw, atol=1e-11 * max(1.0, np.abs(xw).max()))


def test_lu_return_value_follows_the_reference():
    """Factor_unblocked returns (last zero pivot j) + 1, which reads as 0 when
    j = 0 (LU.hpp:63-74); the blocked Factor adds the panel offset and 1 more
    (LU.hpp:94-96, 112-113).  LAPACK's getrf returns (first zero pivot) + 1."""
    rng = np.random.default_rng(9)
    a = _rand(rng, 20, 20, False)
    a[:, 0] = 0.0
    assert O.lu_factor(a)[2] == 0 and lapack.dgetrf(a.copy())[2] == 1
    a = _rand(rng, 20, 20, False)
    a[:, 6] = 0.0       # an exact zero pivot at j = 6
    _, _, info = O.lu_factor(a)
    assert info == 7
    a = _rand(rng, 100, 100, False)
    a[:, 70] = 0.0      # an exact zero pivot at j = 70, in the panel starting at 64
    _, _, info = O.lu_factor(a)
    assert info == 72 and lapack.dgetrf(a.copy())[2] == 71


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_lu_refine_pins(trans, cplx):
    """RefineSolution: berr equals its closed form on the returned x, the
    refined x is at working accuracy, ferr bounds the true forward error,
    and ferr is twice LAPACK gesvx's (the same Hager/Higham estimate of the
    same weighted inverse, with RNP's eps = 2^-52 in the weights against
    LAPACK's 2^-53)."""
    if cplx and trans == "T":
        pytest.skip("LAPACK gesvx trans='T' on complex is covered by 'C'")
    rng = np.random.default_rng(13)
    for n in (8, 120):
        a = _rand(rng, n, n, cplx)
        b = _rand(rng, n, 3, cplx)
        lu, piv, _ = O.lu_factor(a)
        x0 = O.lu_solve(trans, lu, piv, b)
        x, ferr, berr = O.lu_refine(trans, a, lu, piv, b, x0.copy())
        op = a if trans == "N" else (a.T if trans == "T" else a.conj().T)
        n1 = lambda z: np.abs(z.real) + np.abs(z.imag)  # noqa: E731
        for j in range(3):
            r = b[:, j] - op @ x[:, j]
            den = n1(op) @ n1(x[:, j]) + n1(b[:, j])
            assert berr[j] <= 4 * np.finfo(float).eps
            assert abs(berr[j] - np.max(n1(r) / den)) <= 2 * np.finfo(float).eps
        xt = np.linalg.solve(op, b)
        fwd = np.max(np.abs(x - xt), axis=0) / np.max(np.abs(x), axis=0)
        assert np.all(fwd <= ferr)
        res = (lapack.zgesvx if cplx else lapack.dgesvx)(a, b, fact="N", trans=trans)
        # real: the same estimate to 2%; complex: zgerfs' refinement stops at a
        # different x, which moves the estimate by ~10%
        np.testing.assert_allclose(ferr, 2 * res[9], rtol=0.25 if cplx else 0.02)
