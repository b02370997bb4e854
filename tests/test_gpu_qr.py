"""GPU parity of the drop-in QR (Factor / MultQ / GenerateQ) against the oracle.

Every case prints its name and shape before touching the GPU.  Tolerances:
R and tau column-scaled elementwise error <= 1e-11 against the oracle,
||A - QR|| / (n eps ||A||) <= 10 and ||Q^H Q - I|| / (n eps) <= 10.
"""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import orth_error, qr_residual, rel_err_colscaled

pytestmark = pytest.mark.gpu

TOL_ELEM = 1e-11


def rnp_input(m, n, cplx=False, dist=O.UNIFORM_11):
    """The reference harness' input: GenerateVector column by column from the
    internal static seed (tests/test_qr_tall.cpp:13-15)."""
    O.reset_internal_seed()
    return O.generate_matrix(dist, m, n, None, cplx)


def seeded(m, n, cplx, seed):
    iseed = np.array([seed % 4096, (seed * 7) % 4096, (seed * 13) % 4096, (2 * seed + 1) % 4096], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed, cplx)


def check_factor(a, tag):
    print(f"[case] {tag} shape={a.shape} dtype={a.dtype}", flush=True)
    m, n = a.shape
    k = min(m, n)
    af_o, tau_o = O.factor(a)
    af_g, tau_g = R.factor(a)
    r_o, r_g = np.triu(af_o[:k]), np.triu(af_g[:k])
    e_r = rel_err_colscaled(r_g, r_o)
    e_t = float(np.max(np.abs(tau_g - tau_o))) if k else 0.0
    e_v = float(np.max(np.abs(np.tril(af_g, -1) - np.tril(af_o, -1)))) if k else 0.0
    out = {"R": e_r, "tau": e_t, "V": e_v}
    if m >= n and n > 0:
        q = R.generateq(af_g, tau_g, n=n)
        out["res"] = qr_residual(a, q, r_g)
        out["orth"] = orth_error(q)
    elif k > 0:
        q = R.generateq(af_g[:, :m], tau_g, n=m)
        out["res"] = qr_residual(a, q, r_g)
        out["orth"] = orth_error(q)
    print(f"        {out}", flush=True)
    assert e_r <= TOL_ELEM, out
    assert e_t <= TOL_ELEM, out
    if k:
        assert out["res"] <= 10 and out["orth"] <= 10, out
    return af_g, tau_g


@pytest.mark.parametrize("cplx", [False, True])
def test_config1_512x256(cplx):
    a = rnp_input(512, 256, cplx)
    check_factor(a, f"config1 512x256 cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(170, 160), (166, 170)])
def test_reference_test_shapes(shape, cplx):
    a = rnp_input(*shape, cplx)
    check_factor(a, f"ref-test {shape} cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (7, 1), (2, 2), (33, 31), (64, 64), (65, 65), (100, 300),
                                   (1000, 64), (777, 129), (300, 1000), (1500, 900), (2048, 2048)])
def test_shapes(shape, cplx):
    a = seeded(*shape, cplx, seed=sum(shape))
    check_factor(a, f"shape {shape} cplx={cplx}")


def _edge_matrix(cplx, rank_deficient):
    rng = np.random.default_rng(5)
    m, n = 300, 96
    dt = np.complex128 if cplx else np.float64
    a = rng.standard_normal((m, n)).astype(dt)
    if cplx:
        a += 1j * rng.standard_normal((m, n))
    a[:, 3] = 0.0                        # x = 0 and alpha = 0: tau = 0 (H = I)
    a[:, 5] = 0.0
    a[5, 5] = 2.5j if cplx else 2.5      # a single entry (purely imaginary alpha for z)
    a[:, 10] *= 1e-300                   # |x|^2 underflows to 0: the scaled-norm path
    a[:, 11] *= 1e-310                   # subnormal column: |beta| < 2^-971, the rescale loop
    a[:, 12] *= 1e300                    # |x|^2 overflows
    a[:, 13] *= 1e-160                   # tiny but representable squares
    if rank_deficient:
        a[:, 20] = a[:, 19]
    return a


@pytest.mark.parametrize("cplx", [False, True])
def test_edge_columns(cplx):
    """Zero columns (tau = 0), a lone (imaginary) alpha, tiny / subnormal /
    huge columns: R, V and tau match the oracle elementwise everywhere."""
    check_factor(_edge_matrix(cplx, False), f"edge columns cplx={cplx}")


@pytest.mark.parametrize("cplx", [False, True])
def test_rank_deficient(cplx):
    """A repeated column: the reflector after the deficiency is fixed only by
    rounding noise (in any implementation), so parity there is normwise:
    residual and orthogonality, plus elementwise R up to the deficient column."""
    a = _edge_matrix(cplx, True)
    print(f"[case] rank deficient cplx={cplx}", flush=True)
    af_o, tau_o = O.factor(a)
    af_g, tau_g = R.factor(a)
    n = a.shape[1]
    r_g = np.triu(af_g[:n])
    assert rel_err_colscaled(r_g[:, :20], np.triu(af_o[:n])[:, :20]) <= TOL_ELEM
    assert abs(r_g[20, 20]) <= 1e-12 * np.max(np.abs(r_g[:, 19]))
    q = R.generateq(af_g, tau_g, n=n)
    assert qr_residual(a, q, r_g) <= 10
    assert orth_error(q) <= 10


@pytest.mark.parametrize("cplx", [False, True])
def test_nonfinite_input(cplx):
    """NaN / Inf in the input: the call returns normally and the NaN
    propagates into R (no hang, no error code)."""
    a = seeded(200, 50, cplx, 3)
    a[7, 4] = np.nan
    a[9, 30] = np.inf
    print(f"[case] nonfinite cplx={cplx}", flush=True)
    af, tau = R.factor(a)
    assert not np.all(np.isfinite(af))


@pytest.mark.parametrize("cplx", [False, True])
def test_sign_rule_alpha_zero(cplx):
    """alpha = 0, x = [3, 4]: RNP gives beta = +5, v = -x/5, tau = 1
    (Reflector.hpp:131-132), where LAPACK would give beta = -5."""
    a = np.zeros((3, 1), dtype=np.complex128 if cplx else np.float64)
    a[1, 0], a[2, 0] = 3.0, 4.0
    af, tau = R.factor(a)
    assert af[0, 0] == pytest.approx(5.0, rel=1e-15)
    assert tau[0] == pytest.approx(1.0, rel=1e-15)
    assert af[1, 0] == pytest.approx(-0.6, rel=1e-15)
    assert af[2, 0] == pytest.approx(-0.8, rel=1e-15)


def test_zero_and_identity():
    for a in (np.zeros((50, 40)), np.eye(60, 45), np.eye(40, 60)):
        af, tau = check_factor(a, f"structured {a.shape}")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "C"), ("L", "N"), ("R", "N"), ("R", "C")])
def test_multq(side, trans, cplx):
    m, n = 400, 150
    a = seeded(m, n, cplx, 11)
    af, tau = O.factor(a)
    rows, cols = (m, 97) if side == "L" else (97, m)
    c = seeded(rows, cols, cplx, 12)
    print(f"[case] multq {side}{trans} cplx={cplx} C={c.shape}", flush=True)
    got = R.multq(side, trans, af, tau, c)
    ref = O.multq(side, trans, af, tau, c)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print(f"        err={err:.3e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,k", [(300, 300, 300), (500, 200, 200), (500, 200, 150), (129, 70, 3)])
def test_generateq(m, n, k, cplx):
    a = seeded(m, max(n, k), cplx, 21)
    af, tau = O.factor(a)
    print(f"[case] generateq m={m} n={n} k={k} cplx={cplx}", flush=True)
    got = R.generateq(af, tau[:k], n=n, k=k)
    ref = O.generateq(af, tau[:k], n=n, k=k)
    err = float(np.max(np.abs(got - ref)))
    print(f"        err={err:.3e}", flush=True)
    assert err < 1e-13


def test_reference_test_identities():
    """The checks of tests/test_qr_tall.cpp:45-216 on 170x160, d and z, in
    units of m*n*eps (the reference prints 0.1-3.0 for these)."""
    for cplx in (False, True):
        m, n = 170, 160
        a = rnp_input(m, n, cplx)
        print(f"[case] test_qr_tall identities cplx={cplx}", flush=True)
        scale = 1.0 / (m * n * np.finfo(float).eps)
        af, tau = R.factor(a)
        b = R.multq("L", "C", af, tau, a)
        e1 = np.abs(np.triu(af[:n]) - np.triu(b[:n])).sum() * scale
        rr = np.zeros_like(a)
        rr[:n] = np.triu(af[:n])
        b = R.multq("L", "N", af, tau, rr)
        e2 = np.abs(a - b).sum() * scale
        bt = R.multq("R", "N", af, tau, a.conj().T)
        e3 = np.abs(np.triu(af[:n]) - bt.conj().T[:n]).sum() * scale
        bt = R.multq("R", "C", af, tau, rr.conj().T)
        e4 = np.abs(a - bt.conj().T).sum() * scale
        q = R.generateq(af, tau, n=n)
        e5 = np.abs(q.conj().T @ q - np.eye(n)).sum() * scale
        e6 = np.abs(q @ np.triu(af[:n]) - a).sum() * scale
        qf = R.multq("L", "N", af, tau, np.eye(m, dtype=a.dtype))
        e7 = np.abs(a.conj().T @ qf[:, n:]).sum() * scale
        errs = [e1, e2, e3, e4, e5, e6, e7]
        print(f"        {['%.3f' % e for e in errs]}", flush=True)
        assert max(errs) < 30.0


@pytest.mark.parametrize("cplx", [False, True])
def test_untuned_defaults(tmp_path, cplx):
    """With no factorization entry in the tuning table the built-in widths
    apply (64-column real panels in 128-column blocks, 32-column complex
    CholeskyQR2 panels); they match the oracle like the tuned ones."""
    p = tmp_path / "batched_only.json"
    p.write_text('{"version": 1, "entries": [{"kind": "batched", "dtype": "d", "lds": 20480, "blocked_min": 33}]}')
    assert R.tuning_load(str(p)) == 1
    try:
        assert R.tuning_lookup_factor("z" if cplx else "d", 2048, 1024) is None
        a = seeded(2048, 1024, cplx, seed=77)
        check_factor(a, f"untuned 2048x1024 cplx={cplx}")
    finally:
        R.lib().rnpx_tuning_load(None)


def test_profile_families():
    """Live kernel timing selects kernel families: with only the trailing
    update instrumented, no W = V^H C launch is recorded, and enabling resets
    the records."""
    import torch
    m = n = 1024
    print(f"[case] profile families {m}x{n}", flush=True)
    h = R.Handle()
    a = R.colmajor_empty(m, n, torch.float64)
    tau = torch.empty(n, dtype=torch.float64, device="cuda")
    h.larnv(2, [1, 2, 3, 5], a)
    h.profile(True, kernels=["update"])
    h.geqrf(a, tau)
    h.synchronize()
    upd, wg = h.profile_read("update"), h.profile_read("wgemm")
    assert upd["launches"] > 0 and upd["ms"] > 0 and upd["flops"] > 0
    assert wg["launches"] == 0
    h.larnv(2, [1, 2, 3, 5], a)
    h.profile(True)
    h.geqrf(a, tau)
    h.synchronize()
    assert h.profile_read("wgemm")["launches"] > 0
    assert h.profile_read("update")["launches"] == upd["launches"]
    h.profile(False)
    assert h.profile_read("update")["launches"] == 0


@pytest.mark.parametrize("shape", [(4096, 4096), (8192, 2048)])
def test_factor_deterministic(shape):
    """Every reduction on the path sums in a fixed order, so repeated
    factorizations of one matrix on one handle (look-ahead streams, split-K
    products, the block kernel's grid reductions) agree bit for bit."""
    import torch
    m, n = shape
    print(f"[case] deterministic {shape}", flush=True)
    h = R.Handle()
    a0 = R.colmajor_empty(m, n, torch.float64)
    h.larnv(2, [5, 3, 2, 9], a0)
    h.synchronize()  # larnv runs on the handle's stream, the copies below on torch's
    a = R.colmajor_empty(m, n, torch.float64)
    tau = torch.empty(min(m, n), dtype=torch.float64, device="cuda")
    outs = []
    for _ in range(3):
        a.copy_(a0)
        torch.cuda.synchronize()
        h.geqrf(a, tau)
        h.synchronize()
        outs.append((a.clone(), tau.clone()))
    for af, tf in outs[1:]:
        assert torch.equal(af, outs[0][0]) and torch.equal(tf, outs[0][1])
    assert bool(torch.isfinite(outs[0][1]).all())
