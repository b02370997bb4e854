"""The hand-written MFMA products (rnpx_dgemm.hip real, rnpx_zgemm.hip
complex) against numpy: op(A) in N/T/C times B for the alpha/beta forms the
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
