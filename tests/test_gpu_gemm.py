"""The hand-written MFMA products (rnpx_dgemm.hip real, rnpx_zgemm.hip
complex) and the few-column kernel (rnpx_gemv.hip) against numpy: op(A) in N/T/C times B for the alpha/beta forms the
library takes (alpha = +-1 with beta = 1 into the accumulators, beta = 0,
general), ragged shapes (tails in every dimension, a block of reflectors as
M), and the split-K path through the block-reflector apply (W = V^H C split
over the rows, summed in a fixed order).  Tolerance: max |error| <= 1e-13
times the magnitude scale k * max|a| * max|b| (the products are FP64)."""
import numpy as np
import pytest
import torch

import rnp2_amd as R

pytestmark = pytest.mark.gpu

SHAPES = [(128, 96, 64), (257, 131, 67), (33, 200, 300), (1000, 517, 128), (20, 700, 1500), (64, 64, 16)]
FORMS = [(1.0, 0.0), (-1.0, 1.0), (1.0, 1.0), (0.5 - 0.25j, 2.0 + 1.0j), (-2.0, 0.0)]


def dev(x):
    t = R.colmajor_empty(x.shape[0], x.shape[1], torch.complex128 if np.iscomplexobj(x) else torch.float64)
    t.copy_(torch.from_numpy(np.ascontiguousarray(x)).to("cuda"))
    torch.cuda.synchronize()  # the handle's stream reads it next
    return t


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
def test_gemm_forms(cplx, ta):
    rng = np.random.default_rng(7)
    h = R.Handle()
    try:
        for (m, n, k) in SHAPES:
            for alpha, beta in FORMS:
                if not cplx:
                    alpha, beta = complex(alpha).real, complex(beta).real
                def rnd(r, c):
                    x = rng.standard_normal((r, c))
                    return x + 1j * rng.standard_normal((r, c)) if cplx else x
                a = rnd(m, k) if ta == "N" else rnd(k, m)
                b, c = rnd(k, n), rnd(m, n)
                print(f"[case] cplx={cplx} op={ta} {m}x{n}x{k} alpha={alpha} beta={beta}", flush=True)
                A, B, C = dev(a), dev(b), dev(c)
                h.gemm(ta, "N", alpha, A, B, beta, C)
                h.synchronize()
                opa = a if ta == "N" else (a.T if ta == "T" else a.conj().T)
                want = alpha * (opa @ b) + beta * c
                got = C.cpu().numpy()
                scale = k * np.abs(a).max() * np.abs(b).max() + np.abs(c).max()
                assert np.abs(got - want).max() <= 1e-13 * scale
    finally:
        h.close()


GEMV_SHAPES = [(1, 1, 1), (300, 1, 5000), (16384, 1, 2048), (2048, 4, 16384), (517, 3, 67), (70, 2, 0),
               (4099, 4, 31), (256, 1, 256)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
def test_gemv_forms(cplx, ta):
    """One to four columns of B: the HBM-bound kernel with its reduction
    split in slices (row blocks x K chunks for N, columns x row chunks for
    T/C) added in a fixed order, K = 0 included (C = beta C)."""
    rng = np.random.default_rng(11)
    h = R.Handle()
    try:
        for (m, n, k) in GEMV_SHAPES:
            for alpha, beta in FORMS:
                if not cplx:
                    alpha, beta = complex(alpha).real, complex(beta).real
                def rnd(r, c):
                    x = rng.standard_normal((r, c))
                    return x + 1j * rng.standard_normal((r, c)) if cplx else x
                a = rnd(m, k) if ta == "N" else rnd(k, m)
                b, c = rnd(k, n), rnd(m, n)
                print(f"[case] gemv cplx={cplx} op={ta} {m}x{n}x{k} alpha={alpha} beta={beta}", flush=True)
                A = dev(a) if a.size else R.colmajor_empty(a.shape[0], a.shape[1], torch.complex128 if cplx else torch.float64)
                B = dev(b) if b.size else R.colmajor_empty(b.shape[0], b.shape[1], torch.complex128 if cplx else torch.float64)
                C = dev(c)
                h.gemm(ta, "N", alpha, A, B, beta, C)
                h.synchronize()
                opa = a if ta == "N" else (a.T if ta == "T" else a.conj().T)
                want = alpha * (opa @ b) + beta * c
                got = C.cpu().numpy()
                amax = np.abs(a).max() if a.size else 0.0
                bmax = np.abs(b).max() if b.size else 0.0
                scale = k * amax * bmax + np.abs(c).max()
                assert np.abs(got - want).max() <= 1e-13 * scale
            # beta = 0 does not read C: NaN in C is cleared
            c = np.full((m, n), np.nan) * (1 + 1j if cplx else 1)
            a = rnd(m, k) if ta == "N" else rnd(k, m)
            b = rnd(k, n)
            if k == 0:
                continue
            A, B, C = dev(a), dev(b), dev(c)
            h.gemm(ta, "N", 1.0, A, B, 0.0, C)
            h.synchronize()
            opa = a if ta == "N" else (a.T if ta == "T" else a.conj().T)
            assert np.isfinite(C.cpu().numpy()).all()
            assert np.abs(C.cpu().numpy() - opa @ b).max() <= 1e-13 * k * np.abs(a).max() * np.abs(b).max()
    finally:
        h.close()


@pytest.mark.parametrize("cplx", [False, True])
def test_block_reflector_apply_split_k(cplx):
    """C <- H^H C with H = I - V T V^H: the W = V^H C product is split over
    the rows (thousands of them) and the update C -= V Z takes C into the
    accumulators."""
    rng = np.random.default_rng(3)
    h = R.Handle()
    try:
        for (m, n, k) in [(4096, 700, 64), (3000, 1300, 128), (6000, 97, 32)]:
            def rnd(r, c):
                x = rng.standard_normal((r, c))
                return x + 1j * rng.standard_normal((r, c)) if cplx else x
            v, t, c = rnd(m, k), np.triu(rnd(k, k)) * 0.1, rnd(m, n)
            v[:k] = np.tril(v[:k], -1) + np.eye(k)  # the stored form: unit diagonal, zeros above
            print(f"[case] larfb cplx={cplx} {m}x{n} k={k}", flush=True)
            V, T, C = dev(v), dev(t), dev(c)
            h.larfb("L", "C", V, T, C)
            h.synchronize()
            want = c - v @ (t.conj().T @ (v.conj().T @ c))
            got = C.cpu().numpy()
            scale = m * np.abs(v).max() ** 2 * np.abs(c).max() * max(1.0, k * np.abs(t).max())
            assert np.abs(got - want).max() <= 1e-13 * scale
    finally:
        h.close()
