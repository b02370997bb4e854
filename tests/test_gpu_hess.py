"""GPU parity of the callers that reuse the QR kernels:

* Hessenberg::MultQ and Hessenberg::GenerateQ (RNP/LA/Hessenberg.hpp:478-588)
  on LAPACK gehrd reductions, against the oracle's restatement;
* the generalized eigensolver's pre-QR (GeneralizedEigensystems.hpp:459-468:
  QR::Factor of B's block, QR::MultQ("L","C") of it onto A's block, VL set to
  the identity and QR::GenerateQ on its block), run through the drop-in C ABI
  on submatrices addressed by pointer offset and leading dimension, against
  the oracle doing the same steps.

Bar: <= 1e-11 relative to the largest entry, as for MultQ/GenerateQ.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import oracle as O
import rnp2_amd as R

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _rand(rng, m, n, cplx):
    x = rng.standard_normal((m, n))
    if cplx:
        x = x + 1j * rng.standard_normal((m, n))
    return np.asfortranarray(x)


def _gehrd(rng, n, ilo, ihi, cplx):
    a = _rand(rng, n, n, cplx)
    for j in range(n):
        for i in range(j + 1, n):
            if j < ilo or i >= ihi:
                a[i, j] = 0.0
    ht, tau, info = (lapack.zgehrd if cplx else lapack.dgehrd)(a, lo=ilo, hi=max(ihi - 1, 0))
    assert info == 0
    return np.asfortranarray(ht), tau


def _err(got, want):
    return float(np.max(np.abs(got - want)) / max(np.abs(want).max(), 1e-300)) if want.size else 0.0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n,ilo,ihi", [(12, 2, 9), (200, 0, 200), (300, 17, 260), (1000, 0, 1000), (5, 4, 5)])
def test_hess_generateq(n, ilo, ihi, cplx):
    print(f"[case] hess generateq n={n} ilo={ilo} ihi={ihi} cplx={cplx}", flush=True)
    ht, tau = _gehrd(np.random.default_rng(n + ilo), n, ilo, ihi, cplx)
    want = O.hess_generateq(ht.copy(), tau, ilo, ihi)
    got = R.hess_generateq(ht.copy(), tau, ilo, ihi)
    e = _err(got, want)
    print(f"        {e:.2e}", flush=True)
    assert e <= TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "C"), ("R", "N"), ("R", "C")])
def test_hess_multq(side, trans, cplx):
    print(f"[case] hess multq {side}{trans} cplx={cplx}", flush=True)
    rng = np.random.default_rng(7)
    for n, ilo, ihi in ((12, 2, 9), (300, 17, 260)):
        ht, tau = _gehrd(rng, n, ilo, ihi, cplx)
        c = _rand(rng, n, 33, cplx) if side == "L" else _rand(rng, 33, n, cplx)
        want = O.hess_multq(side, trans, ht, tau, c, ilo, ihi)
        got = R.hess_multq(side, trans, ht, tau, c, ilo, ihi)
        e = _err(got, want)
        print(f"        n={n}: {e:.2e}", flush=True)
        assert e <= TOL


def _at(x, i, j):
    """Address of x[i, j] in a Fortran-ordered array (a submatrix pointer)."""
    return ctypes.c_void_p(x.ctypes.data + (i + j * x.shape[0]) * x.itemsize)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n,ilo,ihi", [(40, 0, 40), (300, 11, 270)])
def test_generalized_eigen_pre_qr_through_abi(n, ilo, ihi, cplx):
    """GeneralizedEigensystems.hpp:459-468 with both eigenvector sets wanted:
    irows = ihi-ilo, icols = n-ilo."""
    print(f"[case] gen-eig pre-QR n={n} ilo={ilo} ihi={ihi} cplx={cplx}", flush=True)
    rng = np.random.default_rng(n)
    a = _rand(rng, n, n, cplx)
    b = _rand(rng, n, n, cplx)
    irows, icols = ihi - ilo, n - ilo
    L = R.lib()
    pre = "z" if cplx else "d"
    fac = getattr(L, f"rnpx_{pre}geqrf")
    mq = getattr(L, "rnpx_zunmqr" if cplx else "rnpx_dormqr")
    gq = getattr(L, "rnpx_zungqr" if cplx else "rnpx_dorgqr")
    # device library, through the ABI on submatrices of the full arrays
    ag, bg = a.copy(order="F"), b.copy(order="F")
    tau = np.zeros(n, dtype=a.dtype)
    work = np.zeros(64 * n, dtype=a.dtype)
    lw = ctypes.c_size_t(work.size)
    wp = work.ctypes.data_as(ctypes.c_void_p)
    tp = tau.ctypes.data_as(ctypes.c_void_p)
    assert fac(irows, icols, _at(bg, ilo, ilo), n, tp, ctypes.byref(lw), wp) == 0
    assert mq(b"L", b"C", irows, icols, irows, _at(bg, ilo, ilo), n, tp, _at(ag, ilo, ilo), n, ctypes.byref(lw),
              wp) == 0
    vl = np.asfortranarray(np.eye(n, dtype=a.dtype))
    if irows > 1:
        blk = bg[ilo + 1:ihi, ilo:ihi - 1]
        vl[ilo + 1:ihi, ilo:ihi - 1] = np.tril(blk)
    assert gq(irows, irows, irows, _at(vl, ilo, ilo), n, tp, ctypes.byref(lw), wp) == 0
    # the oracle, the same steps on copies of the blocks
    ao, bo = a.copy(order="F"), b.copy(order="F")
    bf, to = O.factor(bo[ilo:ihi, ilo:])
    bo[ilo:ihi, ilo:] = bf
    ao[ilo:ihi, ilo:] = O.multq("L", "C", bf, to, ao[ilo:ihi, ilo:], k=irows)
    vo = np.eye(n, dtype=a.dtype)
    if irows > 1:
        vo[ilo + 1:ihi, ilo:ihi - 1] = np.tril(bf[1:, :irows - 1])
    vo[ilo:ihi, ilo:ihi] = O.generateq(vo[ilo:ihi, ilo:ihi].copy(order="F"), to, k=irows)
    eb, ea, ev = _err(bg, bo), _err(ag, ao), _err(vl, vo)
    print(f"        B {eb:.2e}  A {ea:.2e}  VL {ev:.2e}", flush=True)
    assert eb <= TOL and ea <= TOL and ev <= TOL
    # the pre-QR's invariants: B's block upper triangular after Factor, VL unitary,
    # VL^H (original) B's block = R
    assert np.allclose(vl.conj().T @ vl, np.eye(n), atol=1e-12)
    rb = np.triu(bg[ilo:ihi, ilo:])
    np.testing.assert_allclose(vl[ilo:ihi, ilo:ihi].conj().T @ b[ilo:ihi, ilo:], rb, atol=1e-10 * np.abs(b).max())
