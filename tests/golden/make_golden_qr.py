"""Writes tests/golden/rnp_reference_qr.npz from the real RNP reference's
blocked QR (RNP/LA/QR.hpp Factor, MultQ, GenerateQ) built with its own
src/BLAS_mix.cpp against the image's MKL (oracle/_ref/librnpref_qr.so,
`make -C oracle ref`).  The values are pinned at RNP + MKL rounding.

    python tests/golden/make_golden_qr.py

Contents (every input that is not regenerable from RNP's own generator is
stored beside the outputs, so the fixture is pure data):
  c1_*        config 1: 512 x 256 real, RNP Uniform_11 stream from the
              default seed (regenerable bit for bit with the oracle's
              generator, itself pinned against the reference's): Afac, tau, Q
  reft_*      the reference tests' shapes 170 x 160 (d and z) and 166 x 170 (d),
              from the same stream: Afac, tau
  edge_*      small edge matrices (alpha = 0, zero / tiny / huge / subnormal
              columns, rank deficiency, zero trailing rows), d and z: input,
              Afac, tau
  mq_*        MultQ for the four side/trans pairs on a 60 x 40 factor
  short_*     a 200 x 140 Factor with a short workspace (nb degrades to 8);
              the input is seeded(200, 140, 999)
  batch_*     problems drawn from the batched configuration (m log-uniform
              in [16, 256], n uniform in [8, m], m n <= 16384) plus one 200 x 150: shapes,
              Afac, tau; problem i's input is seeded(m, n, i)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import refqr  # noqa: E402

OUT = os.path.join(HERE, "rnp_reference_qr.npz")


def rnp_stream(m, n, cplx):
    O.reset_internal_seed()
    return O.generate_matrix(O.UNIFORM_11, m, n, None, cplx)


def edge_cases(cplx):
    rng = np.random.default_rng(11 + cplx)
    dt = np.complex128 if cplx else np.float64

    def g(m, n):
        x = rng.standard_normal((m, n))
        if cplx:
            x = x + 1j * rng.standard_normal((m, n))
        return np.asfortranarray(x.astype(dt))

    out = []
    a = g(40, 20); a[0, 0] = 0.0; out.append(a)                                    # alpha == 0 (sign rule)
    a = g(40, 20); a[:, 3] = 0.0; out.append(a)                                    # zero column: tau = 0 in a block
    a = g(40, 20); a[:, 5] *= 1e-300; a[:, 6] *= 1e-310; out.append(a)             # tiny / subnormal columns
    a = g(40, 20); a[:, 2] *= 1e300; out.append(a)                                 # huge column
    a = g(60, 30); a[:, 10:] = a[:, :20] @ g(20, 20)[:, :20]; out.append(a)        # rank deficient
    a = np.asfortranarray(np.eye(50, 35, dtype=dt)); out.append(a)                 # identity
    a = g(80, 40); a[50:, :] = 0.0; out.append(a)                                  # zero trailing rows
    a = g(1, 1); out.append(a)
    a = g(1, 9); out.append(a)
    a = g(9, 1); out.append(a)
    if cplx:
        a = g(30, 12); a[0, 0] = 2.0j; a[1:, 0] = 0.0; out.append(a)               # lone imaginary alpha
    return out


def seeded(m, n, p):
    """RNP Normal01 from iseed derived from p (regenerable with the oracle's
    generator, which is pinned bit for bit against the reference's)."""
    iseed = np.array([p % 4096, (7 * p + 1) % 4096, (13 * p + 2) % 4096, (2 * p + 1) % 4096], dtype=np.int32)
    return O.generate_matrix(O.NORMAL01, m, n, iseed)


def batch_shapes(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        m = int(np.rint(np.exp(rng.uniform(np.log(16), np.log(256)))))
        n = int(rng.integers(8, m + 1))
        if m * n <= 16384:  # keeps the fixture small
            out.append((m, n))
    # one problem past the crossover (k > 128) so the blocked path runs
    out.append((200, 150))
    return out


def main():
    if not refqr.available():
        raise SystemExit(f"reference QR unavailable: {refqr.why_unavailable()}")
    refqr.set_threads(1)
    d = {}
    a = rnp_stream(512, 256, False)
    af, tau = refqr.factor(a)
    d["c1_afac"], d["c1_tau"] = af, tau
    d["c1_q"] = refqr.generateq(af, tau, n=256)
    for (m, n, cplx) in [(170, 160, False), (166, 170, False), (170, 160, True)]:
        if True:
            a = rnp_stream(m, n, cplx)
            af, tau = refqr.factor(a)
            key = f"reft_{'z' if cplx else 'd'}_{m}x{n}"
            d[key + "_afac"], d[key + "_tau"] = af, tau
    for cplx in (False, True):
        pre = "z" if cplx else "d"
        cases = edge_cases(cplx)
        d[f"edge_{pre}_count"] = np.array(len(cases))
        for i, a in enumerate(cases):
            af, tau = refqr.factor(a)
            d[f"edge_{pre}_{i}_a"], d[f"edge_{pre}_{i}_afac"], d[f"edge_{pre}_{i}_tau"] = a, af, tau
        rng = np.random.default_rng(21 + cplx)
        a = rng.standard_normal((60, 40)) + (1j * rng.standard_normal((60, 40)) if cplx else 0)
        af, tau = refqr.factor(a)
        d[f"mq_{pre}_a"], d[f"mq_{pre}_afac"], d[f"mq_{pre}_tau"] = a, af, tau
        for side in "LR":
            for trans in "NC":
                shape = (60, 30) if side == "L" else (30, 60)
                c = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
                d[f"mq_{pre}_{side}{trans}_c"] = c
                d[f"mq_{pre}_{side}{trans}_out"] = refqr.multq(side, trans, af, tau, c)
    a = seeded(200, 140, 999)
    af, tau = refqr.factor(a, lwork_cap=140 * 8)
    d["short_lwork"], d["short_afac"], d["short_tau"] = np.array(140 * 8), af, tau
    shapes = batch_shapes(10, 41)
    d["batch_shapes"] = np.array(shapes)
    for i, (m, n) in enumerate(shapes):
        af, tau = refqr.factor(seeded(m, n, i))
        d[f"batch_{i}_afac"], d[f"batch_{i}_tau"] = af, tau
    d["blas"] = np.array(refqr.blas_name())
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e6:.2f} MB, {len(d)} arrays)")


if __name__ == "__main__":
    main()
