"""The C ABI and the host-side surface, on the CPU (no kernel is launched).

librnpx.so must load here, export every function include/rnpx.h declares,
answer RNP's workspace queries exactly as the reference does
(RNP/LA/QR.hpp:141-144, 355-358, 519-522), reject bad arguments LAPACK
style, and report RNPX_ERR_NOGPU (never a CPU fallback) when no GPU is
present.  The C++ surface include/rnpx_qr.hpp is compiled and run the same
way.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import rnp2_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rnpx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rnpx_\w+)\s*\(", text)))


def has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40
    L = ctypes.CDLL(R.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared in include/rnpx.h but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(rnpx_\w+)$", out, flags=re.M))
    assert set(names) <= exported


def test_python_signatures_cover_header():
    names = set(declared_functions())
    assert names <= set(R.SIGNATURES), sorted(names - set(R.SIGNATURES))


def test_version_string():
    assert R.version().startswith("rnpx")


def test_workspace_queries_match_reference():
    L = R.lib()
    sz = ctypes.c_size_t
    for m, n in ((512, 256), (8192, 8192), (100, 300)):
        lw = sz(0)
        assert L.rnpx_dgeqrf(m, n, None, m, None, ctypes.byref(lw), None) == 0
        assert lw.value == 32 * n
        lw = sz(0)
        assert L.rnpx_zgeqrf(m, n, None, m, None, ctypes.byref(lw), None) == 0
        assert lw.value == 32 * n
    lw = sz(0)
    assert L.rnpx_dormqr(b"L", b"C", 100, 40, 30, None, 100, None, None, 100, ctypes.byref(lw), None) == 0
    assert lw.value == 32 * 40 + 32 * 32
    lw = sz(0)
    assert L.rnpx_zunmqr(b"R", b"N", 40, 100, 30, None, 100, None, None, 40, ctypes.byref(lw), None) == 0
    assert lw.value == 32 * 40 + 32 * 32
    lw = sz(0)
    assert L.rnpx_dorgqr(100, 40, 30, None, 100, None, ctypes.byref(lw), None) == 0
    assert lw.value == 40 * 32
    # zero sizes return before the query, as in the reference
    lw = sz(7)
    assert L.rnpx_dgeqrf(0, 5, None, 1, None, ctypes.byref(lw), None) == 0 and lw.value == 7


def test_argument_errors_lapack_style():
    L = R.lib()
    sz = ctypes.c_size_t
    lw = sz(0)
    assert L.rnpx_dgeqrf(10, 5, None, 9, None, ctypes.byref(lw), None) == -4       # lda < m
    # 0 < lwork < m is not an argument error: the reference asserts lwork >= m
    # but prints, continues and factors (its own query returns nb n)
    lw = sz(5)
    a = np.zeros(50)
    t = np.zeros(5)
    st = L.rnpx_dgeqrf(10, 5, a.ctypes.data_as(ctypes.c_void_p), 10, t.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(lw), a.ctypes.data_as(ctypes.c_void_p))
    assert st in (0, R.RNPX_ERR_NOGPU)
    assert L.rnpx_dormqr(b"X", b"N", 10, 5, 3, None, 10, None, None, 10, ctypes.byref(sz(0)), None) == -1
    assert L.rnpx_dormqr(b"L", b"Q", 10, 5, 3, None, 10, None, None, 10, ctypes.byref(sz(0)), None) == -2
    assert L.rnpx_dormqr(b"L", b"N", 10, 5, 11, None, 10, None, None, 10, ctypes.byref(sz(0)), None) == -5
    assert L.rnpx_dorgqr(10, 5, 6, None, 10, None, ctypes.byref(sz(0)), None) == -3   # k > n


@pytest.mark.skipif(has_gpu(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    L = R.lib()
    h = ctypes.c_void_p()
    assert L.rnpx_create(ctypes.byref(h), 0) == R.RNPX_ERR_NOGPU
    a = np.asfortranarray(np.random.default_rng(0).standard_normal((64, 32)))
    a0 = a.copy()
    tau = np.zeros(32)
    lw = ctypes.c_size_t(32 * 32)
    work = np.zeros(32 * 32)
    st = L.rnpx_dgeqrf(64, 32, a.ctypes.data_as(ctypes.c_void_p), 64, tau.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(lw), work.ctypes.data_as(ctypes.c_void_p))
    assert st == R.RNPX_ERR_NOGPU
    np.testing.assert_array_equal(a, a0)          # nothing computed on the CPU
    assert R.Factor(64, 32, a, 64, tau, np.array([32 * 32], dtype=np.uint64), work) == R.RNPX_ERR_NOGPU
    with pytest.raises(R.RnpxError):
        R.Handle()


def test_descriptor_layout_matches_c():
    """rnp2_amd.QR_DESC mirrors struct rnpx_qr_desc byte for byte."""
    src = r'''
    #include <stdio.h>
    #include <stddef.h>
    #include "rnpx.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu %zu\n", sizeof(rnpx_qr_desc), offsetof(rnpx_qr_desc, nrhs),
             offsetof(rnpx_qr_desc, a_off), offsetof(rnpx_qr_desc, tau_off), offsetof(rnpx_qr_desc, b_off),
             sizeof(rnpx_request_info));
      return 0;
    }'''
    got = _run_c(src, "gcc", "c")
    size, o_nrhs, o_a, o_tau, o_b, rinfo = map(int, got.split())
    d = R.QR_DESC
    assert d.itemsize == size
    assert d.fields["nrhs"][1] == o_nrhs and d.fields["a_off"][1] == o_a
    assert d.fields["tau_off"][1] == o_tau and d.fields["b_off"][1] == o_b
    assert ctypes.sizeof(R.RequestInfo) == rinfo


def test_make_batch_descriptors():
    shapes = [(20, 10), (7, 7), (256, 9)]
    desc, a_len, tau_len, b_len = R.make_batch(shapes, nrhs=2)
    assert list(desc["m"]) == [20, 7, 256] and list(desc["n"]) == [10, 7, 9]
    assert np.all(desc["lda"] >= desc["m"])
    offs = desc["a_off"]
    for i in range(len(shapes) - 1):            # packed without overlap
        assert offs[i] + desc["lda"][i] * desc["n"][i] <= offs[i + 1]
    assert a_len >= offs[-1] + desc["lda"][-1] * desc["n"][-1]
    assert tau_len == sum(min(m, n) for m, n in shapes)
    assert b_len >= sum(m * 2 for m, n in shapes)


def _run_c(src, cc, ext, extra=()):
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t." + ext)
        open(p, "w").write(src)
        exe = os.path.join(d, "t")
        cmd = [cc, p, "-I", os.path.join(ROOT, "include"), "-o", exe, *extra]
        subprocess.run(cmd, check=True, capture_output=True, text=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60,
                           env={**os.environ, "LD_LIBRARY_PATH": os.path.dirname(R.LIB_PATH)})
        return r.stdout + r.stderr


def test_cpp_surface_compiles_and_queries():
    """include/rnpx_qr.hpp: the reference's template signatures, queries and
    assertion messages, for double, complex<double>, float."""
    src = r'''
    #include <complex>
    #include <cstdio>
    #include "rnpx_qr.hpp"
    using namespace RNP::LA;
    int main() {
      size_t lw = 0;
      QR::Factor<double>(512, 256, NULL, 512, NULL, &lw, NULL);
      std::printf("factor %zu\n", lw);
      lw = 0;
      QR::Factor<std::complex<double> >(100, 300, NULL, 100, NULL, &lw, NULL);
      std::printf("zfactor %zu\n", lw);
      lw = 0;
      QR::MultQ<double>("L", "C", 100, 40, 30, NULL, 100, NULL, NULL, 100, &lw, NULL);
      std::printf("multq %zu\n", lw);
      lw = 0;
      QR::GenerateQ<float>(100, 40, 30, NULL, 100, NULL, &lw, NULL);
      std::printf("genq %zu\n", lw);
      QR::Factor<double>(10, 5, NULL, 9, NULL, &lw, NULL);        /* lda < m */
      double a[50] = {1}, tau[5], work[100];
      lw = 100;
      QR::Factor(10, 5, a, 10, tau, &lw, work);
      std::printf("status %d\n", QR::LastStatus());
      return 0;
    }'''
    out = _run_c(src, "g++", "cpp", ("-std=c++11", "-L", os.path.dirname(R.LIB_PATH), "-lrnpx",
                                     f"-Wl,-rpath,{os.path.dirname(R.LIB_PATH)}"))
    assert "factor 8192" in out and "zfactor 9600" in out
    assert "multq 2304" in out and "genq 1280" in out
    assert "Assertion failed in" in out and "lda >= m" in out
    if not has_gpu():
        assert f"status {R.RNPX_ERR_NOGPU}" in out and "rnpx: Factor: status" in out


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_reads_committed_traffic():
    """The roofline's traffic field comes from the committed counter summary
    of the 8192^2 workload: a per-launch byte count at least the algorithmic
    one (8 (2 mc nc + mc k + k nc) averaged over the launches, 3.69e8 B)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    t = bench.update_traffic_bytes()
    assert t is not None and 3.69e8 <= t <= 3.0 * 3.69e8, t


def test_cpp_lq_surface_compiles_and_queries():
    """include/rnpx_lq.hpp: LQ's template signatures and workspace queries
    (LQ.hpp:135-138 nb*n, 329-333 nb*nw + nb^2, 500-504 n*nb)."""
    src = r'''
    #include <complex>
    #include <cstdio>
    #include "rnpx_lq.hpp"
    using namespace RNP::LA;
    int main() {
      size_t lw = 0;
      LQ::Factor<double>(50, 70, NULL, 50, NULL, &lw, NULL);
      std::printf("factor %zu\n", lw);
      lw = 0;
      LQ::Factor<std::complex<double> >(0, 70, NULL, 1, NULL, &lw, NULL);
      std::printf("empty %zu\n", lw);
      lw = 0;
      LQ::MultQ<double>("R", "N", 40, 100, 30, NULL, 30, NULL, NULL, 40, &lw, NULL);
      std::printf("multq %zu\n", lw);
      lw = 0;
      LQ::GenerateQ<float>(30, 100, 30, NULL, 30, NULL, &lw, NULL);
      std::printf("genq %zu\n", lw);
      LQ::Factor<double>(10, 5, NULL, 9, NULL, &lw, NULL);        /* lda < m */
      double a[50] = {1}, tau[5], work[400];
      lw = 400;
      LQ::Factor(5, 10, a, 5, tau, &lw, work);
      std::printf("status %d\n", LQ::LastStatus());
      return 0;
    }'''
    out = _run_c(src, "g++", "cpp", ("-std=c++11", "-L", os.path.dirname(R.LIB_PATH), "-lrnpx",
                                     f"-Wl,-rpath,{os.path.dirname(R.LIB_PATH)}"))
    assert "factor 2240" in out and "empty 0" in out
    assert "multq 2304" in out and "genq 3200" in out
    assert "Assertion failed in" in out and "lda >= m" in out
    if not has_gpu():
        assert f"status {R.RNPX_ERR_NOGPU}" in out and "rnpx: LQ::Factor: status" in out


def test_ql_rq_workspace_queries_match_reference():
    """QL::Factor nb*n (QL.hpp:139-141), RQ::Factor nb*m (RQ.hpp:142-144),
    MultQ nb*nw + nb^2 (QL.hpp:351-354, RQ.hpp:359-362), QL::GenerateQ n*nb
    (QL.hpp:509-512), RQ::GenerateQ m*nb (RQ.hpp:528-531)."""
    L = R.lib()
    sz = ctypes.c_size_t
    for fn, m, n, want in (("rnpx_dgeqlf", 100, 300, 32 * 300), ("rnpx_zgeqlf", 300, 100, 32 * 100),
                           ("rnpx_dgerqf", 100, 300, 32 * 100), ("rnpx_zgerqf", 300, 100, 32 * 300)):
        lw = sz(0)
        assert getattr(L, fn)(m, n, None, m, None, ctypes.byref(lw), None) == 0
        assert lw.value == want, fn
    for fn in ("rnpx_dormql", "rnpx_zunmql", "rnpx_dormrq", "rnpx_zunmrq"):
        lw = sz(0)
        assert getattr(L, fn)(b"L", b"C", 100, 40, 30, None, 100, None, None, 100, ctypes.byref(lw), None) == 0
        assert lw.value == 32 * 40 + 32 * 32, fn
    lw = sz(0)
    assert L.rnpx_dorgql(100, 40, 30, None, 100, None, ctypes.byref(lw), None) == 0 and lw.value == 40 * 32
    lw = sz(0)
    assert L.rnpx_zungrq(40, 100, 30, None, 40, None, ctypes.byref(lw), None) == 0 and lw.value == 40 * 32
    # argument errors, LAPACK style
    assert L.rnpx_dgeqlf(10, 5, None, 9, None, ctypes.byref(sz(0)), None) == -4
    a = np.zeros(50)
    t = np.zeros(5)
    st = L.rnpx_dgerqf(10, 5, a.ctypes.data_as(ctypes.c_void_p), 10, t.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(sz(5)), a.ctypes.data_as(ctypes.c_void_p))  # 0 < lwork < m: accepted
    assert st in (0, R.RNPX_ERR_NOGPU)
    assert L.rnpx_dormql(b"L", b"N", 10, 5, 3, None, 2, None, None, 10, ctypes.byref(sz(0)), None) == -7
    assert L.rnpx_dorgql(10, 5, 6, None, 10, None, ctypes.byref(sz(0)), None) == -3   # k > n
    assert L.rnpx_dorgrq(5, 10, 6, None, 5, None, ctypes.byref(sz(0)), None) == -3   # k > m


def test_cpp_ql_rq_surfaces_compile_and_query():
    """include/rnpx_ql.hpp and include/rnpx_rq.hpp: the reference's template
    signatures and workspace queries."""
    src = r'''
    #include <complex>
    #include <cstdio>
    #include "rnpx_ql.hpp"
    #include "rnpx_rq.hpp"
    using namespace RNP::LA;
    int main() {
      size_t lw = 0;
      QL::Factor<double>(50, 70, NULL, 50, NULL, &lw, NULL);
      std::printf("qlfactor %zu\n", lw);
      lw = 0;
      RQ::Factor<std::complex<double> >(50, 70, NULL, 50, NULL, &lw, NULL);
      std::printf("rqfactor %zu\n", lw);
      lw = 0;
      QL::MultQ<double>("L", "N", 100, 40, 30, NULL, 100, NULL, NULL, 100, &lw, NULL);
      std::printf("qlmultq %zu\n", lw);
      lw = 0;
      RQ::MultQ<float>("R", "C", 40, 100, 30, NULL, 30, NULL, NULL, 40, &lw, NULL);
      std::printf("rqmultq %zu\n", lw);
      lw = 0;
      QL::GenerateQ<double>(100, 40, 30, NULL, 100, NULL, &lw, NULL);
      std::printf("qlgenq %zu\n", lw);
      lw = 0;
      RQ::GenerateQ<double>(30, 100, 30, NULL, 30, NULL, &lw, NULL);
      std::printf("rqgenq %zu\n", lw);
      RQ::Factor<double>(10, 5, NULL, 9, NULL, &lw, NULL);        /* lda < m */
      double a[50] = {1}, tau[5], work[400];
      lw = 400;
      QL::Factor(10, 5, a, 10, tau, &lw, work);
      std::printf("status %d\n", QL::LastStatus());
      return 0;
    }'''
    out = _run_c(src, "g++", "cpp", ("-std=c++11", "-L", os.path.dirname(R.LIB_PATH), "-lrnpx",
                                     f"-Wl,-rpath,{os.path.dirname(R.LIB_PATH)}"))
    assert "qlfactor 2240" in out and "rqfactor 1600" in out
    assert "qlmultq 2304" in out and "rqmultq 2304" in out
    assert "qlgenq 1280" in out and "rqgenq 960" in out
    assert "Assertion failed in" in out and "lda >= m" in out
    if not has_gpu():
        assert f"status {R.RNPX_ERR_NOGPU}" in out and "rnpx: QL::Factor: status" in out


def test_hessenberg_q_workspace_queries_and_errors():
    """Hessenberg::GenerateQ's query nh*nb (Hessenberg.hpp:553-556) and
    MultQ's pass-through of QR::MultQ's (nb*nw + nb^2; none when nh = 0)."""
    L = R.lib()
    sz = ctypes.c_size_t
    lw = sz(0)
    assert L.rnpx_dorghr(10, 2, 8, None, 10, None, ctypes.byref(lw), None) == 0 and lw.value == 5 * 32
    lw = sz(0)
    assert L.rnpx_zunghr(10, 0, 10, None, 10, None, ctypes.byref(lw), None) == 0 and lw.value == 9 * 32
    lw = sz(0)
    assert L.rnpx_dormhr(b"L", b"N", 10, 5, 2, 8, None, 10, None, None, 10, ctypes.byref(lw), None) == 0
    assert lw.value == 32 * 5 + 32 * 32
    lw = sz(3)
    assert L.rnpx_zunmhr(b"R", b"C", 6, 10, 4, 5, None, 10, None, None, 6, ctypes.byref(sz(0)), None) == 0
    assert L.rnpx_dorghr(10, 8, 8, None, 10, None, ctypes.byref(sz(0)), None) == -3        # ilo == ihi
    assert L.rnpx_dorghr(10, 2, 11, None, 10, None, ctypes.byref(sz(0)), None) == -3       # ihi > n
    assert L.rnpx_dormhr(b"X", b"N", 10, 5, 2, 8, None, 10, None, None, 10, ctypes.byref(sz(0)), None) == -1


def test_lu_argument_errors_and_empty():
    L = R.lib()
    piv = np.zeros(4, dtype=np.uint64)
    pp = piv.ctypes.data_as(ctypes.c_void_p)
    assert L.rnpx_dgetrf(10, 5, None, 9, pp) == -4                    # lda < m
    assert L.rnpx_zgetrf(0, 5, None, 1, pp) == 0                      # empty: nothing to do
    assert L.rnpx_dgetrs(b"X", 4, 1, None, 4, pp, None, 4) == -1       # bad trans
    assert L.rnpx_dgetrs(b"N", 0, 1, None, 1, pp, None, 1) == 0        # empty
    assert L.rnpx_dgetrs(b"N", 4, 1, None, 3, pp, None, 4) == -5       # lda < n
    bad = np.array([0, 9, 2, 3], dtype=np.uint64)
    assert L.rnpx_dgetrs(b"N", 4, 1, None, 4, bad.ctypes.data_as(ctypes.c_void_p), None, 4) == -6  # pivot out of range


def test_getri_query_and_arguments():
    """LU::Invert's workspace query (LU.hpp:149-155): n = 0 returns before it,
    then 64 n; invalid arguments are RNPX_ERR_ARG (-i would read as the
    reference's -(i+1) zero-pivot result)."""
    L = R.lib()
    lw = ctypes.c_size_t(9)
    assert L.rnpx_dgetri(0, None, 1, None, ctypes.byref(lw), None) == 0 and lw.value == 9
    lw = ctypes.c_size_t(0)
    assert L.rnpx_dgetri(100, None, 100, None, ctypes.byref(lw), None) == 0 and lw.value == 6400
    lw = ctypes.c_size_t(0)
    assert L.rnpx_zgetri(30, None, 30, None, ctypes.byref(lw), None) == 0 and lw.value == 64 * 30
    a = np.zeros(16)
    piv = np.zeros(4, dtype=np.uint64)
    lw = ctypes.c_size_t(4)
    assert L.rnpx_dgetri(4, a.ctypes.data_as(ctypes.c_void_p), 3, piv.ctypes.data_as(ctypes.c_void_p),
                         ctypes.byref(lw), a.ctypes.data_as(ctypes.c_void_p)) == R.RNPX_ERR_ARG
    piv[2] = 9
    assert L.rnpx_dgetri(4, a.ctypes.data_as(ctypes.c_void_p), 4, piv.ctypes.data_as(ctypes.c_void_p),
                         ctypes.byref(lw), a.ctypes.data_as(ctypes.c_void_p)) == R.RNPX_ERR_ARG
