"""Writes tests/golden/rnp_reference_vectors.npz from the real RNP reference.

The vectors come from oracle/_ref/librnpref.so, the header-only part of the
reference (RNP/LA/Reflector.hpp Generate, RNP/BLAS/BLAS_mix.hpp Norm2,
RNP/Types.hpp hypot3, RNP/Random.hpp generators, RNP/MatrixGen.hpp
GenerateDiagonal) compiled from its own headers by `make -C oracle ref`.
Run it in a fresh process (the default-seed stream is drawn first):

    python tests/golden/make_golden.py

Inputs are drawn here with numpy and stored beside the reference's outputs,
so the fixture is pure data: tests feed the same inputs to the oracle and,
on a GPU, to the HIP reflector kernels.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "librnpref.so")
OUT = os.path.join(HERE, "rnp_reference_vectors.npz")

_sz, _p = ctypes.c_size_t, ctypes.c_void_p


def load_ref():
    if not os.path.exists(REF_LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    L = ctypes.CDLL(REF_LIB)
    for pre in ("d", "z"):
        getattr(L, f"ref_{pre}_generate").argtypes = [_sz, _p, _p, _sz, _p]
        getattr(L, f"ref_{pre}_norm2").argtypes = [_sz, _p, _sz]
        getattr(L, f"ref_{pre}_norm2").restype = ctypes.c_double
        getattr(L, f"ref_{pre}_generate_vector").argtypes = [ctypes.c_int, _sz, _p, _p]
    L.ref_hypot3.argtypes = [ctypes.c_double] * 3
    L.ref_hypot3.restype = ctypes.c_double
    L.ref_uniform_real_vector.argtypes = [_sz, _p, _p]
    L.ref_generate_diagonal.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _sz, _p, _p]
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def reflector_cases(rng):
    """(alpha, x) inputs covering the branches of Reflector::Generate."""
    cases = []
    for n in (1, 2, 3, 5, 17, 40, 100):
        for cplx in (False, True):
            x = rng.standard_normal(n - 1) + (1j * rng.standard_normal(n - 1) if cplx else 0)
            al = rng.standard_normal() + (1j * rng.standard_normal() if cplx else 0)
            cases.append((al, x))
    special = [
        (0.0, [3.0, 4.0]),                 # alpha == 0: beta positive (sign rule)
        (-0.0, [3.0, 4.0]),
        (2.0, [0.0, 0.0, 0.0]),            # x == 0, real alpha: tau = 0
        (-2.0, []),                        # n == 1
        (1.0, [1e-300, 2e-300]),
        (1e-300, [1e-300, 3e-300]),        # |beta| < 2^-971: rescale branch
        (0.0, [1e-310, 0.0, 4e-310]),      # subnormal column
        (5e-324, [0.0, 5e-324]),
        (1e300, [1e300, -1e300]),          # near overflow
        (-3.0, [1e-200, 1e200]),
        (1.0, [1e-170] * 9),
    ]
    for al, x in special:
        cases.append((al, np.array(x, dtype=np.float64)))
    cspecial = [
        (1j * 2.0, [0.0, 0.0]),            # purely imaginary alpha, x == 0: tau != 0
        (1j * 2.0, [1.0 + 1.0j, -2.0]),    # purely imaginary alpha
        (0.0, [3.0 + 0j, 4.0j]),
        (1e-300 + 1e-300j, [1e-300j, 2e-300]),
        (-1e300 + 1e300j, [1e300, 1e299j]),
        (2.0 + 0j, [0.0, 0.0]),
    ]
    for al, x in cspecial:
        cases.append((complex(al), np.array(x, dtype=np.complex128)))
    return cases


def main():
    L = load_ref()
    rng = np.random.default_rng(20240611)
    out = {}

    # 1) default-seed stream (internal seed, fresh library): first 256 values
    x = np.zeros(256)
    L.ref_uniform_real_vector(256, ptr(x), None)
    out["uniform_default_256"] = x

    # 2) GenerateVector for every distribution, real and complex, seed {1,2,3,5}
    for pre, dt in (("d", np.float64), ("z", np.complex128)):
        for dist in range(5):
            seed = np.array([1, 2, 3, 5], dtype=np.int32)
            v = np.zeros(300, dtype=dt)
            getattr(L, f"ref_{pre}_generate_vector")(dist, v.size, ptr(v), ptr(seed))
            out[f"genvec_{pre}_{dist}"] = v
            out[f"genvec_{pre}_{dist}_seed_after"] = seed.copy()
    # continuing the stream: two calls of 100 equal one of 200 (seed update)
    seed = np.array([2398, 691, 782, 721], dtype=np.int32)
    v = np.zeros(200)
    L.ref_uniform_real_vector(100, ptr(v), ptr(seed))
    w = np.zeros(100)
    L.ref_uniform_real_vector(100, ptr(w), ptr(seed))
    out["uniform_two_calls"] = np.concatenate([v[:100], w])

    # 3) GenerateDiagonal, every mode
    diags = []
    for mode in range(1, 7):
        for reverse in (0, 1):
            for randsign in (0, 1):
                seed = np.array([4, 3, 2, 1], dtype=np.int32)
                d = np.zeros(17)
                L.ref_generate_diagonal(2 if mode == 6 else 0, mode, reverse, randsign, 1e6, 2.5, d.size, ptr(d),
                                        ptr(seed))
                diags.append(np.concatenate([[mode, reverse, randsign], d]))
    out["diag_rows"] = np.array(diags)

    # 4) Norm2 on scaled vectors, real and complex
    n2_in_d, n2_in_z, n2_d, n2_z = [], [], [], []
    for scale in (1.0, 1e-300, 1e-310, 1e300, 1e-160, 1e160):
        xd = rng.standard_normal(33) * scale
        xd[::7] = 0.0
        n2_in_d.append(xd)
        n2_d.append(L.ref_d_norm2(xd.size, ptr(xd), 1))
        xz = (rng.standard_normal(33) + 1j * rng.standard_normal(33)) * scale
        n2_in_z.append(xz)
        n2_z.append(L.ref_z_norm2(xz.size, ptr(xz), 1))
    out["norm2_in_d"], out["norm2_d"] = np.array(n2_in_d), np.array(n2_d)
    out["norm2_in_z"], out["norm2_z"] = np.array(n2_in_z), np.array(n2_z)
    strided = rng.standard_normal(40)
    out["norm2_strided_in"] = strided
    out["norm2_strided_inc3"] = np.array([L.ref_d_norm2(13, ptr(strided), 3)])

    # 5) hypot3
    triples = np.array([[3.0, 4.0, 12.0], [0.0, 0.0, 0.0], [1e300, 1e300, 1e300], [1e-310, 0.0, 3e-310],
                        [-2.0, 0.0, 0.0], [1.0, -1e-20, 1e20], [5e-324, 5e-324, 0.0]])
    out["hypot3_in"] = triples
    out["hypot3"] = np.array([L.ref_hypot3(*t) for t in triples])

    # 6) Reflector::Generate
    for pre, cplx in (("d", False), ("z", True)):
        offs, al_in, x_in, beta, v_out, tau = [0], [], [], [], [], []
        dt = np.complex128 if cplx else np.float64
        for al, x in reflector_cases(rng):
            if not cplx and (isinstance(al, complex) or x.dtype.kind == "c"):
                continue                   # complex cases only feed the complex reflector
            xa = np.ascontiguousarray(x, dtype=dt)
            ala = np.array([al], dtype=dt)
            t = np.zeros(1, dtype=dt)
            al_in.append(ala[0])
            x_in.append(xa.copy())
            getattr(L, f"ref_{pre}_generate")(xa.size + 1, ptr(ala), ptr(xa), 1, ptr(t))
            beta.append(ala[0])
            v_out.append(xa)
            tau.append(t[0])
            offs.append(offs[-1] + xa.size)
        out[f"gen_{pre}_offsets"] = np.array(offs, dtype=np.int64)
        out[f"gen_{pre}_alpha"] = np.array(al_in, dtype=dt)
        out[f"gen_{pre}_x"] = np.concatenate(x_in).astype(dt)
        out[f"gen_{pre}_beta"] = np.array(beta, dtype=dt)
        out[f"gen_{pre}_v"] = np.concatenate(v_out).astype(dt)
        out[f"gen_{pre}_tau"] = np.array(tau, dtype=dt)

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    sys.exit(main())
