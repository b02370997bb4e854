"""CPU oracle for RNP2's LA::QR path -- TEST INFRASTRUCTURE, not the product.

ctypes bindings to ``oracle/liboracle.so`` (the C restatement in
``rnp_oracle.c`` / ``rnp_oracle_gen.h``; reference citations live there).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker or the timed CPU baseline.

Matrices are numpy arrays in Fortran (column-major) order, float64 or
complex128, mirroring RNP's column-major, 0-based, size_t conventions.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

UNIFORM01, UNIFORM_11, NORMAL01, UNITDISC, UNITCIRCLE = range(5)
(DIAG_GIVEN, DIAG_FIRST1_REMAINING_COND, DIAG_LASTCOND_REMAINING1,
 DIAG_GRADED_EXPONENTIALLY, DIAG_GRADED_LINEARLY, DIAG_RANDOM_EXPONENTIAL,
 DIAG_RANDOM_DIST) = range(7)

_sz = ctypes.c_size_t
_p = ctypes.c_void_p


class Tuning(ctypes.Structure):
    """QR::Tuning<T> values (QR.hpp:41-51): nb, nbmin, crossover nx."""
    _fields_ = [("nb", _sz), ("nbmin", _sz), ("nx", _sz)]


class _Z(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def build():
    """Compile liboracle.so with make (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    for pre in ("d", "z"):
        scal = ctypes.c_double if pre == "d" else _Z
        getattr(L, f"orc_{pre}_norm2").restype = ctypes.c_double
        getattr(L, f"orc_{pre}_norm2").argtypes = [_sz, _p, _sz]
        getattr(L, f"orc_{pre}_generate").argtypes = [_sz, _p, _p, _sz, _p]
        getattr(L, f"orc_{pre}_apply").argtypes = [ctypes.c_char_p, ctypes.c_int, _sz, _sz, _p, _sz,
                                                   scal, _p, _sz, _p]
        getattr(L, f"orc_{pre}_generate_block_tr").argtypes = [_sz, _sz, _p, _sz, _p, _p, _sz]
        getattr(L, f"orc_{pre}_apply_block").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz,
                                                         _sz, _p, _sz, _p, _sz, _p, _sz, _p, _sz]
        getattr(L, f"orc_{pre}_factor_unblocked").argtypes = [_sz, _sz, _p, _sz, _p, _p]
        getattr(L, f"orc_{pre}_factor").argtypes = [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p,
                                                    ctypes.POINTER(Tuning)]
        getattr(L, f"orc_{pre}_multq").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz, _sz,
                                                   _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p,
                                                   ctypes.POINTER(Tuning)]
        getattr(L, f"orc_{pre}_multq_unblocked").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz,
                                                             _sz, _sz, _p, _sz, _p, _p, _sz, _p]
        getattr(L, f"orc_{pre}_generateq").argtypes = [_sz, _sz, _sz, _p, _sz, _p,
                                                       ctypes.POINTER(_sz), _p,
                                                       ctypes.POINTER(Tuning)]
        getattr(L, f"orc_{pre}_generateq_unblocked").argtypes = [_sz, _sz, _sz, _p, _sz, _p, _p]
        getattr(L, f"orc_{pre}_ls_solve").argtypes = [_sz, _sz, _sz, _p, _sz, _p, _sz]
        getattr(L, f"orc_{pre}_lq_factor").argtypes = [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]
        getattr(L, f"orc_{pre}_lq_multq").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz, _sz,
                                                      _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]
        getattr(L, f"orc_{pre}_lq_generateq").argtypes = [_sz, _sz, _sz, _p, _sz, _p,
                                                          ctypes.POINTER(_sz), _p]
        getattr(L, f"orc_{pre}_hess_multq").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz, _sz, _sz,
                                                        _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]
        getattr(L, f"orc_{pre}_hess_generateq").argtypes = [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]
        for nm in ("lu_factor", "lu_factor_unblocked"):
            f = getattr(L, f"orc_{pre}_{nm}")
            f.restype = ctypes.c_int
            f.argtypes = [_sz, _sz, _p, _sz, _p]
        getattr(L, f"orc_{pre}_lu_solve").argtypes = [ctypes.c_char_p, _sz, _sz, _p, _sz, _p, _p, _sz]
        getattr(L, f"orc_{pre}_lu_invert").restype = ctypes.c_int
        getattr(L, f"orc_{pre}_lu_invert").argtypes = [_sz, _p, _sz, _p, _p]
        getattr(L, f"orc_{pre}_lu_refine").argtypes = [ctypes.c_char_p, _sz, _sz, _p, _sz, _p, _sz, _p, _p, _sz,
                                                       _p, _sz, _p, _p]
        for fam in ("ql", "rq"):
            getattr(L, f"orc_{pre}_{fam}_factor").argtypes = [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]
            getattr(L, f"orc_{pre}_{fam}_multq").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz, _sz,
                                                             _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]
            getattr(L, f"orc_{pre}_{fam}_generateq").argtypes = [_sz, _sz, _sz, _p, _sz, _p,
                                                                 ctypes.POINTER(_sz), _p]
        getattr(L, f"orc_{pre}_gemm").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz, _sz,
                                                  scal, _p, _sz, _p, _sz, scal, _p, _sz]
        getattr(L, f"orc_{pre}_generate_vector").argtypes = [ctypes.c_int, _sz, _p, _p]
        for nm in ("last_nonzero_column_limit", "last_nonzero_row_limit"):
            f = getattr(L, f"orc_{pre}_{nm}")
            f.restype = _sz
            f.argtypes = [_sz, _sz, _p, _sz]
    L.orc_hypot3.restype = ctypes.c_double
    L.orc_hypot3.argtypes = [ctypes.c_double] * 3
    L.orc_d_factor_panels.argtypes = [_sz, _sz, _p, _sz, _p, _p, _sz, _sz]
    L.orc_uniform_real_vector.argtypes = [_sz, _p, _p]
    L.orc_reset_internal_seed.argtypes = []
    L.orc_generate_diagonal.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, _sz, _p, _p]
    L.orc_d_random_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, _sz, _sz, _p, _sz, _p]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_get_threads.restype = ctypes.c_int


def _pre(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d"
    if dt == np.complex128:
        return "z"
    raise TypeError(f"oracle supports float64 and complex128, not {dt}")


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def _scal(pre, v):
    if pre == "d":
        return ctypes.c_double(float(np.real(v)))
    v = complex(v)
    return _Z(v.real, v.imag)


def _fortran(a, dtype):
    return np.array(a, dtype=dtype, order="F", copy=True)


def _tuning(tune):
    if tune is None:
        return None
    if isinstance(tune, Tuning):
        return ctypes.byref(tune)
    nb, nbmin, nx = tune
    return ctypes.byref(Tuning(nb, nbmin, nx))


# ---------------------------------------------------------------------------
# Reflectors and level-1
# ---------------------------------------------------------------------------

def norm2(x):
    x = np.ascontiguousarray(x)
    pre = _pre(x.dtype)
    return getattr(lib(), f"orc_{pre}_norm2")(x.size, _ptr(x), 1)


def hypot3(x, y, z):
    return lib().orc_hypot3(x, y, z)


def generate(alpha, x):
    """Reflector::Generate: returns (beta, v, tau) for [alpha; x]."""
    x = np.array(x, copy=True)
    if x.dtype.kind != "c" and isinstance(alpha, complex):
        x = x.astype(np.complex128)
    dtype = np.complex128 if (x.dtype.kind == "c" or isinstance(alpha, complex)) else np.float64
    x = np.ascontiguousarray(x, dtype=dtype)
    pre = _pre(dtype)
    al = np.array([alpha], dtype=dtype)
    tau = np.zeros(1, dtype=dtype)
    getattr(lib(), f"orc_{pre}_generate")(x.size + 1, _ptr(al), _ptr(x), 1, _ptr(tau))
    return al[0], x, tau[0]


def apply(side, vone, v, tau, c):
    """Reflector::Apply (vconj false): returns H C or C H, H = I - tau v v^H."""
    c = _fortran(c, np.result_type(c, v, np.float64))
    pre = _pre(c.dtype)
    v = np.ascontiguousarray(v, dtype=c.dtype)
    m, n = c.shape
    work = np.zeros(max(m, n, 1), dtype=c.dtype)
    getattr(lib(), f"orc_{pre}_apply")(side.encode(), int(vone), m, n, _ptr(v), 1,
                                       _scal(pre, tau), _ptr(c), max(m, 1), _ptr(work))
    return c


def generate_block_tr(v, tau):
    """Reflector::GenerateBlockTr("F","C"): T (k-by-k upper) for V n-by-k."""
    v = _fortran(v, np.result_type(v, np.float64))
    pre = _pre(v.dtype)
    tau = np.ascontiguousarray(tau, dtype=v.dtype)
    n, k = v.shape
    t = np.zeros((k, k), dtype=v.dtype, order="F")
    getattr(lib(), f"orc_{pre}_generate_block_tr")(n, k, _ptr(v), max(n, 1), _ptr(tau),
                                                   _ptr(t), max(k, 1))
    return t


def apply_block(side, trans, v, t, c):
    """Reflector::ApplyBlock(side, trans, "F", "C", ...)."""
    c = _fortran(c, np.result_type(c, v, np.float64))
    pre = _pre(c.dtype)
    v = _fortran(v, c.dtype)
    t = _fortran(t, c.dtype)
    m, n = c.shape
    k = v.shape[1]
    ldwork = n if side[0] == "L" else m
    work = np.zeros((max(ldwork, 1), max(k, 1)), dtype=c.dtype, order="F")
    getattr(lib(), f"orc_{pre}_apply_block")(side.encode(), trans.encode(), m, n, k, _ptr(v),
                                             v.shape[0], _ptr(t), max(t.shape[0], 1), _ptr(c),
                                             max(m, 1), _ptr(work), max(ldwork, 1))
    return c


def last_nonzero_column_limit(a):
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    return getattr(lib(), f"orc_{pre}_last_nonzero_column_limit")(m, n, _ptr(a), max(m, 1))


def last_nonzero_row_limit(a):
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    return getattr(lib(), f"orc_{pre}_last_nonzero_row_limit")(m, n, _ptr(a), max(m, 1))


# ---------------------------------------------------------------------------
# QR
# ---------------------------------------------------------------------------

def factor(a, tune=None, lwork=None, unblocked=False):
    """QR::Factor: returns (afac, tau).  lwork < optimal exercises the
    short-workspace degrade of QR.hpp:155-161."""
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    k = min(m, n)
    tau = np.zeros(max(k, 1), dtype=a.dtype)
    lda = max(m, 1)
    if unblocked:
        work = np.zeros(max(n, 1), dtype=a.dtype)
        getattr(lib(), f"orc_{pre}_factor_unblocked")(m, n, _ptr(a), lda, _ptr(tau), _ptr(work))
        return a, tau[:k]
    lw = _sz(0)
    f = getattr(lib(), f"orc_{pre}_factor")
    f(m, n, _ptr(a), lda, _ptr(tau), ctypes.byref(lw), None, _tuning(tune))
    if lwork is not None:
        lw = _sz(lwork)
    work = np.zeros(max(lw.value, 1), dtype=a.dtype)
    f(m, n, _ptr(a), lda, _ptr(tau), ctypes.byref(lw), _ptr(work), _tuning(tune))
    return a, tau[:k]


def factor_query(m, n, dtype=np.float64, tune=None):
    lw = _sz(0)
    getattr(lib(), f"orc_{_pre(dtype)}_factor")(m, n, None, max(m, 1), None, ctypes.byref(lw),
                                                None, _tuning(tune))
    return lw.value


def factor_panels(a, nb, npanels):
    """The first `npanels` panels of QR::Factor's blocked loop (in place)."""
    assert a.dtype == np.float64 and a.flags.f_contiguous
    m, n = a.shape
    tau = np.zeros(min(m, n), dtype=np.float64)
    work = np.zeros(n * nb, dtype=np.float64)
    lib().orc_d_factor_panels(m, n, _ptr(a), m, _ptr(tau), _ptr(work), nb, npanels)
    return tau


def multq(side, trans, afac, tau, c, k=None, tune=None, lwork=None, unblocked=False):
    """QR::MultQ: returns op(Q) C or C op(Q)."""
    c = _fortran(c, np.result_type(c, afac, np.float64))
    pre = _pre(c.dtype)
    afac = _fortran(afac, c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    if k is None:
        k = tau.size
    lda = afac.shape[0]
    if unblocked:
        work = np.zeros(max(m, n, 1), dtype=c.dtype)
        getattr(lib(), f"orc_{pre}_multq_unblocked")(side.encode(), trans.encode(), m, n, k,
                                                     _ptr(afac), lda, _ptr(tau), _ptr(c),
                                                     max(m, 1), _ptr(work))
        return c
    f = getattr(lib(), f"orc_{pre}_multq")
    lw = _sz(0)
    f(side.encode(), trans.encode(), m, n, k, _ptr(afac), lda, _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), None, _tuning(tune))
    if lwork is not None:
        lw = _sz(lwork)
    work = np.zeros(max(lw.value, 1), dtype=c.dtype)
    f(side.encode(), trans.encode(), m, n, k, _ptr(afac), lda, _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), _ptr(work), _tuning(tune))
    return c


def generateq(afac, tau, n=None, k=None, tune=None, lwork=None, unblocked=False):
    """QR::GenerateQ: overwrite the first n columns of afac by Q."""
    a = _fortran(afac, np.result_type(afac, np.float64))
    pre = _pre(a.dtype)
    m = a.shape[0]
    if n is None:
        n = a.shape[1]
    a = np.asfortranarray(a[:, :n])
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    if k is None:
        k = min(tau.size, n)
    if unblocked:
        work = np.zeros(max(n, 1), dtype=a.dtype)
        getattr(lib(), f"orc_{pre}_generateq_unblocked")(m, n, k, _ptr(a), max(m, 1), _ptr(tau),
                                                         _ptr(work))
        return a
    f = getattr(lib(), f"orc_{pre}_generateq")
    lw = _sz(0)
    f(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None, _tuning(tune))
    if lwork is not None:
        lw = _sz(lwork)
    work = np.zeros(max(lw.value, 1), dtype=a.dtype)
    f(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work), _tuning(tune))
    return a


def lq_factor(a):
    """LQ::Factor (LQ.hpp:126-190): returns (afac, tau); L on and below the
    diagonal, the conjugated reflectors in the rows above it."""
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    k = min(m, n)
    tau = np.zeros(max(k, 1), dtype=a.dtype)
    f = getattr(lib(), f"orc_{pre}_lq_factor")
    lw = _sz(0)
    f(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, 1), dtype=a.dtype)
    f(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work))
    return a, tau[:k]


def lq_factor_query(m, n, dtype=np.float64):
    lw = _sz(0)
    getattr(lib(), f"orc_{_pre(dtype)}_lq_factor")(m, n, None, max(m, 1), None, ctypes.byref(lw), None)
    return lw.value


def lq_multq(side, trans, afac, tau, c, k=None):
    """LQ::MultQ (LQ.hpp:309-405): op(Q) C or C op(Q)."""
    c = _fortran(c, np.result_type(c, afac, np.float64))
    pre = _pre(c.dtype)
    afac = _fortran(afac, c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    if k is None:
        k = tau.size
    lda = afac.shape[0]
    f = getattr(lib(), f"orc_{pre}_lq_multq")
    lw = _sz(0)
    f(side.encode(), trans.encode(), m, n, k, _ptr(afac), lda, _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, n, 1), dtype=c.dtype)
    f(side.encode(), trans.encode(), m, n, k, _ptr(afac), lda, _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), _ptr(work))
    return c


def lq_generateq(afac, tau, m=None, k=None):
    """LQ::GenerateQ (LQ.hpp:491-559): the first m rows of Q (m <= n)."""
    a = _fortran(afac, np.result_type(afac, np.float64))
    pre = _pre(a.dtype)
    n = a.shape[1]
    if m is None:
        m = a.shape[0]
    a = np.asfortranarray(a[:m, :])
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    if k is None:
        k = min(tau.size, m)
    f = getattr(lib(), f"orc_{pre}_lq_generateq")
    lw = _sz(0)
    f(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, 1), dtype=a.dtype)
    f(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work))
    return a


def _fam_factor(fam, a):
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    k = min(m, n)
    tau = np.zeros(max(k, 1), dtype=a.dtype)
    f = getattr(lib(), f"orc_{pre}_{fam}_factor")
    lw = _sz(0)
    f(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, n, 1), dtype=a.dtype)
    f(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work))
    return a, tau[:k]


def _fam_query(fam, what, *args, dtype=np.float64):
    lw = _sz(0)
    f = getattr(lib(), f"orc_{_pre(dtype)}_{fam}_{what}")
    if what == "factor":
        m, n = args
        f(m, n, None, max(m, 1), None, ctypes.byref(lw), None)
    elif what == "multq":
        side, trans, m, n, k = args
        f(side.encode(), trans.encode(), m, n, k, None, max(k, 1), None, None, max(m, 1), ctypes.byref(lw), None)
    else:
        m, n, k = args
        f(m, n, k, None, max(m, 1), None, ctypes.byref(lw), None)
    return lw.value


def _fam_multq(fam, side, trans, afac, tau, c, k=None):
    c = _fortran(c, np.result_type(c, afac, np.float64))
    pre = _pre(c.dtype)
    afac = _fortran(afac, c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    if k is None:
        k = tau.size
    lda = afac.shape[0]
    f = getattr(lib(), f"orc_{pre}_{fam}_multq")
    lw = _sz(0)
    f(side.encode(), trans.encode(), m, n, k, _ptr(afac), lda, _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, n, 1), dtype=c.dtype)
    f(side.encode(), trans.encode(), m, n, k, _ptr(afac), lda, _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), _ptr(work))
    return c


def _fam_generateq(fam, a, tau, k=None):
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    if k is None:
        k = tau.size
    f = getattr(lib(), f"orc_{pre}_{fam}_generateq")
    lw = _sz(0)
    f(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, n, 1), dtype=a.dtype)
    f(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work))
    return a


def ql_factor(a):
    """QL::Factor (QL.hpp:126-205): returns (afac, tau); L in the lower
    triangle ending at the bottom-right corner, the reflectors above it."""
    return _fam_factor("ql", a)


def ql_multq(side, trans, afac, tau, c, k=None):
    """QL::MultQ (QL.hpp:335-430): op(Q) C or C op(Q); afac holds the k
    reflector columns (the last k columns of the factored matrix)."""
    return _fam_multq("ql", side, trans, afac, tau, c, k)


def ql_generateq(a, tau, k=None):
    """QL::GenerateQ (QL.hpp:501-565): a (m x n, m >= n, reflectors in its
    last k columns) becomes the last n columns of Q."""
    return _fam_generateq("ql", a, tau, k)


def rq_factor(a):
    """RQ::Factor (RQ.hpp:131-205): returns (afac, tau); R in the upper
    triangle ending at the bottom-right corner, the conjugated reflectors
    in the rows left of it."""
    return _fam_factor("rq", a)


def rq_multq(side, trans, afac, tau, c, k=None):
    """RQ::MultQ (RQ.hpp:344-440): op(Q) C or C op(Q); afac holds the k
    reflector rows (the last k rows of the factored matrix)."""
    return _fam_multq("rq", side, trans, afac, tau, c, k)


def rq_generateq(a, tau, k=None):
    """RQ::GenerateQ (RQ.hpp:519-580): a (m x n, m <= n, reflectors in its
    last k rows) becomes the last m rows of Q."""
    return _fam_generateq("rq", a, tau, k)


def lu_refine(trans, a, lu, piv, b, x):
    """LU::RefineSolution (LU.hpp:216-366): returns (refined x, ferr, berr)."""
    x = _fortran(x, np.result_type(x, a, np.float64))
    a = _fortran(a, x.dtype)
    lu = _fortran(lu, x.dtype)
    b = _fortran(b, x.dtype)
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    n, nrhs = x.shape
    ferr = np.zeros(max(nrhs, 1))
    berr = np.zeros(max(nrhs, 1))
    getattr(lib(), f"orc_{_pre(x.dtype)}_lu_refine")(trans.encode(), n, nrhs, _ptr(a), a.shape[0], _ptr(lu),
                                                     lu.shape[0], piv.ctypes.data_as(ctypes.c_void_p), _ptr(b),
                                                     b.shape[0], _ptr(x), max(n, 1), _ptr(ferr), _ptr(berr))
    return x, ferr[:nrhs], berr[:nrhs]


def hess_multq(side, trans, afac, tau, c, ilo, ihi):
    """Hessenberg::MultQ (Hessenberg.hpp:478-514): op(Q) C or C op(Q) for the
    Q of a Hessenberg reduction held in afac/tau (0-based ilo, exclusive ihi)."""
    c = _fortran(c, np.result_type(c, afac, np.float64))
    pre = _pre(c.dtype)
    afac = _fortran(afac, c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    f = getattr(lib(), f"orc_{pre}_hess_multq")
    lw = _sz(0)
    f(side.encode(), trans.encode(), m, n, ilo, ihi, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, m, n, 1), dtype=c.dtype)
    lw = _sz(work.size)
    f(side.encode(), trans.encode(), m, n, ilo, ihi, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c), max(m, 1),
      ctypes.byref(lw), _ptr(work))
    return c


def hess_generateq(afac, tau, ilo, ihi):
    """Hessenberg::GenerateQ (Hessenberg.hpp:540-588): the n x n Q."""
    a = _fortran(afac, np.result_type(afac, np.float64))
    pre = _pre(a.dtype)
    n = a.shape[0]
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    f = getattr(lib(), f"orc_{pre}_hess_generateq")
    lw = _sz(0)
    f(n, ilo, ihi, _ptr(a), max(n, 1), _ptr(tau), ctypes.byref(lw), None)
    work = np.zeros(max(lw.value, n, 1), dtype=a.dtype)
    lw = _sz(work.size)
    f(n, ilo, ihi, _ptr(a), max(n, 1), _ptr(tau), ctypes.byref(lw), _ptr(work))
    return a


def hess_query(what, *args, dtype=np.float64):
    lw = _sz(0)
    if what == "generateq":
        n, ilo, ihi = args
        getattr(lib(), f"orc_{_pre(dtype)}_hess_generateq")(n, ilo, ihi, None, max(n, 1), None, ctypes.byref(lw),
                                                             None)
    else:
        side, trans, m, n, ilo, ihi = args
        getattr(lib(), f"orc_{_pre(dtype)}_hess_multq")(side.encode(), trans.encode(), m, n, ilo, ihi, None,
                                                         max(m, n, 1), None, None, max(m, 1), ctypes.byref(lw), None)
    return lw.value


def lu_factor(a, unblocked=False):
    """LU::Factor (LU.hpp:77-115): returns (LU packed, pivots (0-based
    absolute rows, uint64), the reference's return value)."""
    a = _fortran(a, np.result_type(a, np.float64))
    pre = _pre(a.dtype)
    m, n = a.shape
    piv = np.zeros(max(min(m, n), 1), dtype=np.uint64)
    f = getattr(lib(), f"orc_{pre}_lu_factor_unblocked" if unblocked else f"orc_{pre}_lu_factor")
    info = f(m, n, _ptr(a), max(m, 1), piv.ctypes.data_as(ctypes.c_void_p))
    return a, piv[:min(m, n)], info


def lu_invert(lu, piv):
    """LU::Invert (LU.hpp:146-214): returns (inv(A), the reference's return
    value: 0, or -(i+1) for a zero U_ii with the input left unchanged)."""
    a = _fortran(lu, np.result_type(lu, np.float64)).copy(order="F")
    n = a.shape[0]
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    work = np.zeros(max(n, 1), dtype=a.dtype)
    info = getattr(lib(), f"orc_{_pre(a.dtype)}_lu_invert")(n, _ptr(a), max(n, 1), piv.ctypes.data_as(ctypes.c_void_p),
                                                            _ptr(work))
    return a, info


def lu_solve(trans, lu, piv, b):
    """LU::Solve (LU.hpp:117-144): op(A) X = B from the factors."""
    b = _fortran(b, np.result_type(b, lu, np.float64))
    pre = _pre(b.dtype)
    lu = _fortran(lu, b.dtype)
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    n, nrhs = b.shape
    getattr(lib(), f"orc_{pre}_lu_solve")(trans.encode(), n, nrhs, _ptr(lu), lu.shape[0],
                                          piv.ctypes.data_as(ctypes.c_void_p), _ptr(b), max(n, 1))
    return b


def ql_query(what, *args, dtype=np.float64):
    return _fam_query("ql", what, *args, dtype=dtype)


def rq_query(what, *args, dtype=np.float64):
    return _fam_query("rq", what, *args, dtype=dtype)


def ls_solve(a, b):
    """min ||A x - B|| via Factor + MultQ("L","C") + back substitution."""
    a = _fortran(a, np.result_type(a, b, np.float64))
    pre = _pre(a.dtype)
    b2 = b.reshape(b.shape[0], -1)
    b2 = _fortran(b2, a.dtype)
    m, n = a.shape
    getattr(lib(), f"orc_{pre}_ls_solve")(m, n, b2.shape[1], _ptr(a), m, _ptr(b2), m)
    x = b2[:n].copy()
    return x.reshape((n,) + b.shape[1:])


def gemm(transa, transb, alpha, a, b, beta=0.0, c=None):
    dtype = np.result_type(a, b, np.float64)
    pre = _pre(dtype)
    a = _fortran(a, dtype)
    b = _fortran(b, dtype)
    m = a.shape[0] if transa == "N" else a.shape[1]
    k = a.shape[1] if transa == "N" else a.shape[0]
    n = b.shape[1] if transb == "N" else b.shape[0]
    c = np.zeros((m, n), dtype=dtype, order="F") if c is None else _fortran(c, dtype)
    getattr(lib(), f"orc_{pre}_gemm")(transa.encode(), transb.encode(), m, n, k,
                                      _scal(pre, alpha), _ptr(a), max(a.shape[0], 1), _ptr(b),
                                      max(b.shape[0], 1), _scal(pre, beta), _ptr(c), max(m, 1))
    return c


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


# ---------------------------------------------------------------------------
# Generators (harness)
# ---------------------------------------------------------------------------

def uniform_real_vector(n, iseed=None):
    """Random::UniformRealVector; iseed is a 4-int numpy array updated in place,
    or None for the internal static seed."""
    x = np.zeros(n, dtype=np.float64)
    if iseed is not None:
        assert iseed.dtype == np.int32
    lib().orc_uniform_real_vector(n, _ptr(x), _ptr(iseed))
    return x


def reset_internal_seed():
    lib().orc_reset_internal_seed()


def generate_vector(dist, n, iseed=None, complex_=False):
    dtype = np.complex128 if complex_ else np.float64
    x = np.zeros(n, dtype=dtype)
    if iseed is not None:
        assert iseed.dtype == np.int32
    getattr(lib(), f"orc_{_pre(dtype)}_generate_vector")(dist, n, _ptr(x), _ptr(iseed))
    return x


def generate_matrix(dist, m, n, iseed=None, complex_=False):
    """Column-by-column GenerateVector fill, as the reference tests do
    (tests/test_qr_tall.cpp:13-15)."""
    dtype = np.complex128 if complex_ else np.float64
    a = np.zeros((m, n), dtype=dtype, order="F")
    for j in range(n):
        a[:, j] = generate_vector(dist, m, iseed, complex_)
    return a


def generate_diagonal(mode, n, cond, dscale=1.0, dist=UNIFORM01, reverse=False, randsign=False,
                      iseed=None):
    d = np.zeros(n, dtype=np.float64)
    lib().orc_generate_diagonal(dist, mode, int(reverse), int(randsign), cond, dscale, n, _ptr(d),
                                _ptr(iseed))
    return d


def random_matrix(m, n, cond, mode=DIAG_GRADED_EXPONENTIALLY, dmax=1.0, dist=UNIFORM01,
                  iseed=None):
    a = np.zeros((m, n), dtype=np.float64, order="F")
    lib().orc_d_random_matrix(dist, mode, 0, cond, dmax, m, n, _ptr(a), max(m, 1), _ptr(iseed))
    return a
