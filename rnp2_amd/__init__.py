"""rnp2_amd -- MI355X (gfx950) blocked Householder QR, drop-in for RNP2's LA::QR.

Python view of the C ABI in ``include/rnpx.h`` (library ``rnp2_amd/librnpx.so``).
The functions mirror RNP's QR interface (RNP/LA/QR.hpp): ``Factor``,
``MultQ`` and ``GenerateQ`` with the same argument meaning, column-major
arrays, 0-based, and the reference's workspace-query convention; the
lower-case helpers wrap them for numpy arrays and torch device tensors.

There is no CPU fallback: every call goes to the HIP library, and a missing
library or GPU raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librnpx.so")
_lib = None

RNPX_OK = 0
RNPX_ERR_HIP = -1000
RNPX_ERR_NOGPU = -1001
RNPX_ERR_NOMEM = -1002
RNPX_ERR_SYNC = -1003
RNPX_ERR_UNSUPPORTED = -1004
RNPX_ERR_NOTCONV = -1005
RNPX_ERR_ARG = -1006

_sz = ctypes.c_size_t
_p = ctypes.c_void_p
_i = ctypes.c_int
_s = ctypes.c_char_p


class RnpxError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what} failed with status {status}")
        self.status = status


class ZComplex(ctypes.Structure):
    """rnpx_zcomplex (layout of std::complex<double>)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class Tuning(ctypes.Structure):
    """rnpx_tuning: the QR::Tuning<T> hook (QR.hpp:41-51)."""
    _fields_ = [("nb", _i), ("panel_rpt", _i), ("ksplit", _i), ("lookahead", _i), ("panel", _i), ("panel_cus", _i)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", _sz), ("workers", _i), ("mode", _i), ("max_iter", _i), ("tol", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", _i), ("path", _i), ("iters", _i), ("worker", _i), ("berr", ctypes.c_double),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("admitted", ctypes.c_longlong), ("completed", ctypes.c_longlong),
                ("rejected", ctypes.c_longlong), ("failed", ctypes.c_longlong), ("batches", ctypes.c_longlong),
                ("queued", ctypes.c_longlong), ("bytes_in_use", _sz)]


# name -> (restype, argtypes); kept in the order of include/rnpx.h
SIGNATURES = {
    "rnpx_dgeqrf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zgeqrf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dlast_nonzero_column_limit": (_i, [_sz, _sz, _p, _sz, ctypes.POINTER(_sz)]),
    "rnpx_zlast_nonzero_column_limit": (_i, [_sz, _sz, _p, _sz, ctypes.POINTER(_sz)]),
    "rnpx_dlast_nonzero_row_limit": (_i, [_sz, _sz, _p, _sz, ctypes.POINTER(_sz)]),
    "rnpx_zlast_nonzero_row_limit": (_i, [_sz, _sz, _p, _sz, ctypes.POINTER(_sz)]),
    "rnpx_dormqr": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_zunmqr": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_dorgqr": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zungqr": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dgelqf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zgelqf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dormlq": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_zunmlq": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_dorglq": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zunglq": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dgeqlf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dormql": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_dorgql": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dgerqf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dormrq": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_dorgrq": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zgeqlf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zunmql": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_zungql": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zgerqf": (_i, [_sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zunmrq": (_i, [_s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_zungrq": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_create": (_i, [ctypes.POINTER(_p), _sz]),
    "rnpx_destroy": (_i, [_p]),
    "rnpx_set_stream": (_i, [_p, _p]),
    "rnpx_get_stream": (_p, [_p]),
    "rnpx_set_tuning": (_i, [_p, ctypes.POINTER(Tuning)]),
    "rnpx_get_tuning": (_i, [_p, ctypes.POINTER(Tuning)]),
    "rnpx_synchronize": (_i, [_p]),
    "rnpx_workspace_bytes": (_sz, [_p]),
    "rnpx_profile_enable": (_i, [_p, _i]),
    "rnpx_profile_read": (_i, [_p, _i, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "rnpx_exact_path_count": (_i, [_p, ctypes.POINTER(ctypes.c_long)]),
    "rnpx_dgeqrf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_zgeqrf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_dgelqf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_zgelqf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_dgerfs": (_i, [_s, _sz, _sz, _p, _sz, _p, _sz, _p, _p, _sz, _p, _sz, _p, _p, _p, _p]),
    "rnpx_zgerfs": (_i, [_s, _sz, _sz, _p, _sz, _p, _sz, _p, _p, _sz, _p, _sz, _p, _p, _p, _p]),
    "rnpx_dgetrf": (_i, [_sz, _sz, _p, _sz, _p]),
    "rnpx_zgetrf": (_i, [_sz, _sz, _p, _sz, _p]),
    "rnpx_dgetrs": (_i, [_s, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_dgetri": (_i, [_sz, _p, _sz, _p, _p, _p]),
    "rnpx_zgetri": (_i, [_sz, _p, _sz, _p, _p, _p]),
    "rnpx_zgetrs": (_i, [_s, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_dormhr": (_i, [_s, _s, _sz, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_zunmhr": (_i, [_s, _s, _sz, _sz, _sz, _sz, _p, _sz, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "rnpx_dorghr": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_zunghr": (_i, [_sz, _sz, _sz, _p, _sz, _p, ctypes.POINTER(_sz), _p]),
    "rnpx_dorghr_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p]),
    "rnpx_zunghr_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p]),
    "rnpx_dgeqlf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_zgeqlf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_dgerqf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_zgerqf_dev": (_i, [_p, _sz, _sz, _p, _sz, _p]),
    "rnpx_dormqr_dev": (_i, [_p, _s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_zunmqr_dev": (_i, [_p, _s, _s, _sz, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_dorgqr_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p]),
    "rnpx_zungqr_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p]),
    "rnpx_dgels_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_zgels_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_dgels_mixed_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p, _sz, _p, _sz, _i, ctypes.c_double,
                                  ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double)]),
    "rnpx_dls_berr_dev": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p, _sz, _p, _sz, ctypes.POINTER(ctypes.c_double)]),
    "rnpx_dgeqrf_batched": (_i, [_p, _i, _p, _p, _p, _p]),
    "rnpx_dgeqrf_batched_noalloc": (_i, [_p, _i, _p, _p, _p, _p]),
    "rnpx_tuning_load": (_i, [ctypes.c_char_p]),
    "rnpx_tuning_lookup_factor": (_i, [ctypes.c_char, _sz, _sz, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "rnpx_tuning_lookup_batched": (_i, [ctypes.c_char, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "rnpx_tuning_lookup_panel_cus": (_i, [ctypes.c_char, _sz, _sz, ctypes.POINTER(_i)]),
    "rnpx_zgeqrf_batched": (_i, [_p, _i, _p, _p, _p, _p]),
    "rnpx_dgeqrf_batched_plan": (_i, [_p, _i, _p, _p, _p, _p, ctypes.POINTER(_p)]),
    "rnpx_zgeqrf_batched_plan": (_i, [_p, _i, _p, _p, _p, _p, ctypes.POINTER(_p)]),
    "rnpx_batched_plan_run": (_i, [_p]),
    "rnpx_batched_plan_destroy": (_i, [_p]),
    "rnpx_dlarft_dev": (_i, [_p, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_zlarft_dev": (_i, [_p, _sz, _sz, _p, _sz, _p, _p, _sz]),
    "rnpx_dlarfb_dev": (_i, [_p, _s, _s, _sz, _sz, _sz, _p, _sz, _p, _sz, _p, _sz]),
    "rnpx_zlarfb_dev": (_i, [_p, _s, _s, _sz, _sz, _sz, _p, _sz, _p, _sz, _p, _sz]),
    "rnpx_dgemm_dev": (_i, [_p, _s, _s, _sz, _sz, _sz, ctypes.c_double, _p, _sz, _p, _sz, ctypes.c_double, _p,
                            _sz]),
    "rnpx_zgemm_dev": (_i, [_p, _s, _s, _sz, _sz, _sz, None, _p, _sz, _p, _sz, None, _p, _sz]),
    "rnpx_dlarnv_dev": (_i, [_p, _i, ctypes.POINTER(_i), _sz, _p]),
    "rnpx_zlarnv_dev": (_i, [_p, _i, ctypes.POINTER(_i), _sz, _p]),
    "rnpx_server_create": (_i, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_p)]),
    "rnpx_server_submit": (_i, [_p, _sz, _sz, _sz, _p, _sz, _p, _sz, ctypes.POINTER(ctypes.c_longlong)]),
    "rnpx_server_poll": (_i, [_p, ctypes.c_longlong, _i, _p, _sz, ctypes.POINTER(RequestInfo)]),
    "rnpx_server_release": (_i, [_p, ctypes.c_longlong]),
    "rnpx_server_stats_get": (_i, [_p, ctypes.POINTER(ServerStats)]),
    "rnpx_server_destroy": (_i, [_p]),
    "rnpx_version": (_s, []),
    "rnpx_device_name": (_i, [ctypes.c_char_p, _sz]),
}


# rnpx_qr_desc (include/rnpx.h): six int32 then three int64 offsets.
QR_DESC = np.dtype([("m", np.int32), ("n", np.int32), ("lda", np.int32), ("ldb", np.int32), ("nrhs", np.int32),
                    ("pad", np.int32), ("a_off", np.int64), ("tau_off", np.int64), ("b_off", np.int64)])


def make_batch(shapes, nrhs=0):
    """Descriptor table for problems of the given (m, n) shapes packed one after
    another (lda = m, ldb = m).  Returns (desc, a_elems, tau_elems, b_elems)."""
    desc = np.zeros(len(shapes), dtype=QR_DESC)
    a = t = b = 0
    for i, (m, n) in enumerate(shapes):
        r = nrhs if (nrhs and m >= n) else 0
        desc[i] = (m, n, max(m, 1), max(m, 1), r, 0, a, t, b)
        a += m * n
        t += min(m, n)
        b += m * r
    return desc, a, t, b


def lib():
    """Load librnpx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rnp2_amd.build` "
                               "(there is no CPU fallback)")
        # librnpx and torch both need libamdhip64.so.7; whichever loads first
        # serves the whole process, and torch only initialises on its own
        # bundled runtime.  Load torch's first when torch is installed.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = [ZComplex if a is None else a for a in args]
        _lib = L
    return _lib


def _check(status, what):
    if status != RNPX_OK:
        raise RnpxError(status, what)


def _pre(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d"
    if dt == np.complex128:
        return "z"
    raise TypeError(f"rnp2_amd supports float64 and complex128, not {dt}")


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# Mirror of RNP::LA::QR (host numpy arrays in, results in place)
# ---------------------------------------------------------------------------

def Factor(m, n, a, lda, tau, lwork, work):
    """QR::Factor(m, n, a, lda, tau, &lwork, work) -- QR.hpp:129-131.

    ``a``/``tau``/``work`` are numpy arrays (or None for ``work`` in a query);
    ``lwork`` is a one-element numpy uint64 array, updated by a query.
    Returns the status."""
    pre = _pre(a.dtype)
    lw = _sz(int(lwork[0]))
    st = getattr(lib(), f"rnpx_{pre}geqrf")(m, n, _ptr(a), lda, _ptr(tau), ctypes.byref(lw), _ptr(work))
    lwork[0] = lw.value
    return st


def MultQ(side, trans, m, n, k, a, lda, tau, c, ldc, lwork, work):
    """QR::MultQ(side, trans, m, n, k, a, lda, tau, c, ldc, &lwork, work) -- QR.hpp:337-341."""
    pre = _pre(c.dtype)
    fn = "rnpx_dormqr" if pre == "d" else "rnpx_zunmqr"
    lw = _sz(int(lwork[0]))
    st = getattr(lib(), fn)(side.encode(), trans.encode(), m, n, k, _ptr(a), lda, _ptr(tau), _ptr(c), ldc,
                            ctypes.byref(lw), _ptr(work))
    lwork[0] = lw.value
    return st


def GenerateQ(m, n, k, a, lda, tau, lwork, work):
    """QR::GenerateQ(m, n, k, a, lda, tau, &lwork, work) -- QR.hpp:510-512."""
    pre = _pre(a.dtype)
    fn = "rnpx_dorgqr" if pre == "d" else "rnpx_zungqr"
    lw = _sz(int(lwork[0]))
    st = getattr(lib(), fn)(m, n, k, _ptr(a), lda, _ptr(tau), ctypes.byref(lw), _ptr(work))
    lwork[0] = lw.value
    return st


# ---------------------------------------------------------------------------
# Convenience: numpy arrays (column-major copies are made as needed)
# ---------------------------------------------------------------------------

def factor(a):
    """QR-factor a numpy matrix; returns (packed R/V factor, tau)."""
    a = np.array(a, dtype=np.result_type(a, np.float64), order="F", copy=True)
    m, n = a.shape
    k = min(m, n)
    tau = np.zeros(max(k, 1), dtype=a.dtype)
    lw = np.zeros(1, dtype=np.uint64)
    _check(Factor(m, n, a, max(m, 1), tau, lw, None), "Factor query")
    work = np.zeros(max(int(lw[0]), 1), dtype=a.dtype)
    _check(Factor(m, n, a, max(m, 1), tau, lw, work), "Factor")
    return a, tau[:k]


def last_nonzero_limits(a):
    """(LastNonzeroColumnLimit, LastNonzeroRowLimit) of a numpy matrix, on
    the device -- Reflector.hpp:36-75."""
    a = np.asfortranarray(a, dtype=np.result_type(a, np.float64))
    m, n = a.shape
    pre = _pre(a.dtype)
    out = []
    for which in ("column", "row"):
        lim = _sz(0)
        _check(getattr(lib(), f"rnpx_{pre}last_nonzero_{which}_limit")(m, n, _ptr(a), max(m, 1), ctypes.byref(lim)),
               f"last_nonzero_{which}_limit")
        out.append(lim.value)
    return tuple(out)


def exact_path_count(handle=None):
    """Factorizations on the handle (None: the drop-in calls') whose matrix
    held Inf or NaN and so took the reference's blocking and zero trims."""
    c = ctypes.c_long(0)
    _check(lib().rnpx_exact_path_count(handle.h if handle is not None else None, ctypes.byref(c)),
           "exact_path_count")
    return c.value


def multq(side, trans, afac, tau, c, k=None):
    """op(Q) C or C op(Q) for Q held in (afac, tau)."""
    c = np.array(c, dtype=np.result_type(c, afac, np.float64), order="F", copy=True)
    afac = np.asfortranarray(afac, dtype=c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    k = tau.size if k is None else k
    lw = np.zeros(1, dtype=np.uint64)
    _check(MultQ(side, trans, m, n, k, afac, afac.shape[0], tau, c, max(m, 1), lw, None), "MultQ query")
    work = np.zeros(max(int(lw[0]), 1), dtype=c.dtype)
    _check(MultQ(side, trans, m, n, k, afac, afac.shape[0], tau, c, max(m, 1), lw, work), "MultQ")
    return c


def generateq(afac, tau, n=None, k=None):
    """The first n columns of Q from (afac, tau)."""
    a = np.array(afac, dtype=np.result_type(afac, np.float64), order="F", copy=True)
    m = a.shape[0]
    n = a.shape[1] if n is None else n
    a = np.asfortranarray(a[:, :n])
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    k = min(tau.size, n) if k is None else k
    lw = np.zeros(1, dtype=np.uint64)
    _check(GenerateQ(m, n, k, a, max(m, 1), tau, lw, None), "GenerateQ query")
    work = np.zeros(max(int(lw[0]), 1), dtype=a.dtype)
    _check(GenerateQ(m, n, k, a, max(m, 1), tau, lw, work), "GenerateQ")
    return a


def lq_factor(a):
    """LQ-factor a numpy matrix (LQ::Factor): returns (L and conjugated
    reflector rows packed, tau)."""
    a = np.array(a, dtype=np.result_type(a, np.float64), order="F", copy=True)
    m, n = a.shape
    k = min(m, n)
    tau = np.zeros(max(k, 1), dtype=a.dtype)
    pre = _pre(a.dtype)
    lw = _sz(0)
    fn = getattr(lib(), f"rnpx_{pre}gelqf")
    _check(fn(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None), "LQ::Factor query")
    work = np.zeros(max(lw.value, 1), dtype=a.dtype)
    _check(fn(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work)), "LQ::Factor")
    return a, tau[:k]


def lq_multq(side, trans, afac, tau, c, k=None):
    """LQ::MultQ: op(Q) C or C op(Q) for the LQ factor (afac, tau)."""
    c = np.array(c, dtype=np.result_type(c, afac, np.float64), order="F", copy=True)
    afac = np.asfortranarray(afac, dtype=c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    k = tau.size if k is None else k
    pre = _pre(c.dtype)
    fn = getattr(lib(), "rnpx_dormlq" if pre == "d" else "rnpx_zunmlq")
    lw = _sz(0)
    _check(fn(side.encode(), trans.encode(), m, n, k, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c), max(m, 1),
              ctypes.byref(lw), None), "LQ::MultQ query")
    work = np.zeros(max(lw.value, 1), dtype=c.dtype)
    _check(fn(side.encode(), trans.encode(), m, n, k, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c), max(m, 1),
              ctypes.byref(lw), _ptr(work)), "LQ::MultQ")
    return c


def lq_generateq(afac, tau, m=None, k=None):
    """LQ::GenerateQ: the first m rows of Q (m <= n)."""
    a = np.array(afac, dtype=np.result_type(afac, np.float64), order="F", copy=True)
    n = a.shape[1]
    m = a.shape[0] if m is None else m
    a = np.asfortranarray(a[:m, :])
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    k = min(tau.size, m) if k is None else k
    pre = _pre(a.dtype)
    fn = getattr(lib(), "rnpx_dorglq" if pre == "d" else "rnpx_zunglq")
    lw = _sz(0)
    _check(fn(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None), "LQ::GenerateQ query")
    work = np.zeros(max(lw.value, 1), dtype=a.dtype)
    _check(fn(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work)), "LQ::GenerateQ")
    return a


_FAM = {"ql": ("QL", ("geqlf", "ormql", "orgql"), ("geqlf", "unmql", "ungql")),
        "rq": ("RQ", ("gerqf", "ormrq", "orgrq"), ("gerqf", "unmrq", "ungrq"))}


def _fam_fn(fam, which, pre):
    name, dn, zn = _FAM[fam]
    return getattr(lib(), "rnpx_" + pre + (dn if pre == "d" else zn)[which]), name


def _fam_factor(fam, a):
    a = np.array(a, dtype=np.result_type(a, np.float64), order="F", copy=True)
    m, n = a.shape
    k = min(m, n)
    tau = np.zeros(max(k, 1), dtype=a.dtype)
    fn, name = _fam_fn(fam, 0, _pre(a.dtype))
    lw = _sz(0)
    _check(fn(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None), f"{name}::Factor query")
    work = np.zeros(max(lw.value, m, 1), dtype=a.dtype)
    _check(fn(m, n, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work)), f"{name}::Factor")
    return a, tau[:k]


def _fam_multq(fam, side, trans, afac, tau, c, k=None):
    c = np.array(c, dtype=np.result_type(c, afac, np.float64), order="F", copy=True)
    afac = np.asfortranarray(afac, dtype=c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    k = tau.size if k is None else k
    fn, name = _fam_fn(fam, 1, _pre(c.dtype))
    lw = _sz(0)
    _check(fn(side.encode(), trans.encode(), m, n, k, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c), max(m, 1),
              ctypes.byref(lw), None), f"{name}::MultQ query")
    work = np.zeros(max(lw.value, 1), dtype=c.dtype)
    _check(fn(side.encode(), trans.encode(), m, n, k, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c), max(m, 1),
              ctypes.byref(lw), _ptr(work)), f"{name}::MultQ")
    return c


def _fam_generateq(fam, a, tau, k=None):
    a = np.array(a, dtype=np.result_type(a, np.float64), order="F", copy=True)
    m, n = a.shape
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    k = tau.size if k is None else k
    fn, name = _fam_fn(fam, 2, _pre(a.dtype))
    lw = _sz(0)
    _check(fn(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), None), f"{name}::GenerateQ query")
    work = np.zeros(max(lw.value, 1), dtype=a.dtype)
    _check(fn(m, n, k, _ptr(a), max(m, 1), _ptr(tau), ctypes.byref(lw), _ptr(work)), f"{name}::GenerateQ")
    return a


def ql_factor(a):
    """QL-factor a numpy matrix (QL::Factor): returns (L and the reflector
    columns packed, tau)."""
    return _fam_factor("ql", a)


def ql_multq(side, trans, afac, tau, c, k=None):
    """QL::MultQ: op(Q) C or C op(Q); afac holds the k reflector columns."""
    return _fam_multq("ql", side, trans, afac, tau, c, k)


def ql_generateq(a, tau, k=None):
    """QL::GenerateQ: a (m x n, m >= n, reflectors in its last k columns)
    becomes the last n columns of Q."""
    return _fam_generateq("ql", a, tau, k)


def rq_factor(a):
    """RQ-factor a numpy matrix (RQ::Factor): returns (R and the conjugated
    reflector rows packed, tau)."""
    return _fam_factor("rq", a)


def rq_multq(side, trans, afac, tau, c, k=None):
    """RQ::MultQ: op(Q) C or C op(Q); afac holds the k reflector rows."""
    return _fam_multq("rq", side, trans, afac, tau, c, k)


def rq_generateq(a, tau, k=None):
    """RQ::GenerateQ: a (m x n, m <= n, reflectors in its last k rows)
    becomes the last m rows of Q."""
    return _fam_generateq("rq", a, tau, k)


def lu_factor(a):
    """LU::Factor (LU.hpp:77-115) of a numpy matrix on the GPU: returns (LU
    packed, pivots (uint64, 0-based absolute rows), the reference's return
    value)."""
    a = np.array(a, dtype=np.result_type(a, np.float64), order="F", copy=True)
    m, n = a.shape
    piv = np.zeros(max(min(m, n), 1), dtype=np.uint64)
    fn = getattr(lib(), f"rnpx_{_pre(a.dtype)}getrf")
    info = fn(m, n, _ptr(a), max(m, 1), piv.ctypes.data_as(ctypes.c_void_p))
    if info < 0:
        _check(info, "LU::Factor")
    return a, piv[:min(m, n)], info


def lu_solve(trans, lu, piv, b):
    """LU::Solve (LU.hpp:117-144): op(A) X = B from the factors."""
    b = np.array(b, dtype=np.result_type(b, lu, np.float64), order="F", copy=True)
    lu = np.asfortranarray(lu, dtype=b.dtype)
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    n, nrhs = b.shape
    fn = getattr(lib(), f"rnpx_{_pre(b.dtype)}getrs")
    _check(fn(trans.encode(), n, nrhs, _ptr(lu), lu.shape[0], piv.ctypes.data_as(ctypes.c_void_p), _ptr(b),
              max(n, 1)), "LU::Solve")
    return b


def lu_invert(lu, piv):
    """LU::Invert (LU.hpp:146-214) on the GPU: returns (inv(A), the
    reference's return value: 0, or -(i+1) for a zero U_ii with the factors
    returned unchanged)."""
    a = np.array(lu, dtype=np.result_type(lu, np.float64), order="F", copy=True)
    n = a.shape[0]
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    lw = np.zeros(1, dtype=np.uint64)
    fn = getattr(lib(), f"rnpx_{_pre(a.dtype)}getri")
    _check(fn(n, None, max(n, 1), None, lw.ctypes.data_as(ctypes.c_void_p), None), "LU::Invert query")
    work = np.zeros(max(int(lw[0]), 1), dtype=a.dtype)
    info = fn(n, _ptr(a), max(n, 1), piv.ctypes.data_as(ctypes.c_void_p), lw.ctypes.data_as(ctypes.c_void_p),
              _ptr(work))
    if info <= -1000:
        _check(info, "LU::Invert")
    return a, info


def lu_refine(trans, a, lu, piv, b, x):
    """LU::RefineSolution (LU.hpp:216-366) on the GPU: returns (refined x,
    ferr, berr)."""
    x = np.array(x, dtype=np.result_type(x, a, np.float64), order="F", copy=True)
    a = np.asfortranarray(a, dtype=x.dtype)
    lu = np.asfortranarray(lu, dtype=x.dtype)
    b = np.asfortranarray(b, dtype=x.dtype)
    piv = np.ascontiguousarray(piv, dtype=np.uint64)
    n, nrhs = x.shape
    ferr = np.zeros(max(nrhs, 1))
    berr = np.zeros(max(nrhs, 1))
    fn = getattr(lib(), f"rnpx_{_pre(x.dtype)}gerfs")
    _check(fn(trans.encode(), n, nrhs, _ptr(a), a.shape[0], _ptr(lu), lu.shape[0],
              piv.ctypes.data_as(ctypes.c_void_p), _ptr(b), b.shape[0], _ptr(x), max(n, 1), _ptr(ferr), _ptr(berr),
              None, None), "LU::RefineSolution")
    return x, ferr[:nrhs], berr[:nrhs]


def hess_multq(side, trans, afac, tau, c, ilo, ihi):
    """Hessenberg::MultQ: op(Q) C or C op(Q) for the Q of a Hessenberg
    reduction in (afac, tau); 0-based ilo, exclusive ihi."""
    c = np.array(c, dtype=np.result_type(c, afac, np.float64), order="F", copy=True)
    afac = np.asfortranarray(afac, dtype=c.dtype)
    tau = np.ascontiguousarray(tau, dtype=c.dtype)
    m, n = c.shape
    fn = getattr(lib(), "rnpx_dormhr" if _pre(c.dtype) == "d" else "rnpx_zunmhr")
    lw = _sz(0)
    _check(fn(side.encode(), trans.encode(), m, n, ilo, ihi, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c),
              max(m, 1), ctypes.byref(lw), None), "Hessenberg::MultQ query")
    work = np.zeros(max(lw.value, m, n, 1), dtype=c.dtype)
    lw = _sz(work.size)
    _check(fn(side.encode(), trans.encode(), m, n, ilo, ihi, _ptr(afac), afac.shape[0], _ptr(tau), _ptr(c),
              max(m, 1), ctypes.byref(lw), _ptr(work)), "Hessenberg::MultQ")
    return c


def hess_generateq(afac, tau, ilo, ihi):
    """Hessenberg::GenerateQ: the n x n Q of a Hessenberg reduction."""
    a = np.array(afac, dtype=np.result_type(afac, np.float64), order="F", copy=True)
    n = a.shape[0]
    tau = np.ascontiguousarray(tau, dtype=a.dtype)
    if tau.size == 0:
        tau = np.zeros(1, dtype=a.dtype)
    fn = getattr(lib(), "rnpx_dorghr" if _pre(a.dtype) == "d" else "rnpx_zunghr")
    lw = _sz(0)
    _check(fn(n, ilo, ihi, _ptr(a), max(n, 1), _ptr(tau), ctypes.byref(lw), None), "Hessenberg::GenerateQ query")
    work = np.zeros(max(lw.value, n, 1), dtype=a.dtype)
    lw = _sz(work.size)
    _check(fn(n, ilo, ihi, _ptr(a), max(n, 1), _ptr(tau), ctypes.byref(lw), _ptr(work)), "Hessenberg::GenerateQ")
    return a


def tuning_load(path=None):
    """Replace the autotuning table (rnpx_tuning_load); returns the entry count."""
    st = lib().rnpx_tuning_load(None if path is None else str(path).encode())
    if st < 0:
        raise RnpxError(st, "rnpx_tuning_load")
    return st


def tuning_lookup_factor(dtype, m, n):
    """(nb, ksplit) of the table's entry for an m x n factorization, or None."""
    nb, ks = ctypes.c_int(), ctypes.c_int()
    found = lib().rnpx_tuning_lookup_factor(dtype.encode(), m, n, ctypes.byref(nb), ctypes.byref(ks))
    return (nb.value, ks.value) if found == 1 else None


def tuning_lookup_panel_cus(dtype, m, n):
    """panel_cus of the table's entry for an m x n factorization, or None."""
    v = ctypes.c_int()
    found = lib().rnpx_tuning_lookup_panel_cus(dtype.encode(), m, n, ctypes.byref(v))
    return v.value if found == 1 else None


def tuning_lookup_batched(dtype):
    lds, bm = ctypes.c_int(), ctypes.c_int()
    found = lib().rnpx_tuning_lookup_batched(dtype.encode(), ctypes.byref(lds), ctypes.byref(bm))
    return (lds.value, bm.value) if found == 1 else None


def version():
    return lib().rnpx_version().decode()


def device_name():
    buf = ctypes.create_string_buffer(256)
    _check(lib().rnpx_device_name(buf, 256), "device_name")
    return buf.value.decode()


# ---------------------------------------------------------------------------
# Device-resident work through a handle (torch tensors as device memory)
# ---------------------------------------------------------------------------

def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _tpre(t):
    import torch
    if t.dtype == torch.float64:
        return "d"
    if t.dtype == torch.complex128:
        return "z"
    raise TypeError(f"unsupported dtype {t.dtype}")


def _ld(t):
    """Leading dimension of a column-major (Fortran-strided) 2-D tensor."""
    if t.dim() == 1:
        return max(t.shape[0], 1)
    if t.stride(0) != 1:
        raise ValueError("device matrices must be column-major (stride(0) == 1)")
    return max(t.stride(1), t.shape[0], 1)


def colmajor_empty(m, n, dtype, device="cuda"):
    """A column-major m x n device tensor (the transpose of a row-major n x m)."""
    import torch
    return torch.empty((n, m), dtype=dtype, device=device).t()


class Handle:
    """An rnpx handle: a HIP stream plus device workspace."""

    def __init__(self, hbm_budget_bytes=0, stream=None):
        h = ctypes.c_void_p()
        _check(lib().rnpx_create(ctypes.byref(h), hbm_budget_bytes), "rnpx_create")
        self.h = h
        if stream is not None:
            _check(lib().rnpx_set_stream(self.h, ctypes.c_void_p(stream)), "rnpx_set_stream")

    def close(self):
        if self.h:
            lib().rnpx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().rnpx_get_stream(self.h)

    def set_tuning(self, nb=0, panel_rpt=0, ksplit=0, lookahead=0, panel=0, panel_cus=0):
        """panel: 0 automatic, 1 column-by-column reflectors only, 2 CholeskyQR2 +
        Householder reconstruction for real 64-column panels where it applies.
        panel_cus: CUs reserved for the look-ahead panel stream (0 automatic,
        -1 no partition)."""
        t = Tuning(nb, panel_rpt, ksplit, lookahead, panel, panel_cus)
        _check(lib().rnpx_set_tuning(self.h, ctypes.byref(t)), "rnpx_set_tuning")

    KERNELS = {"update": 0, "wgemm": 1, "panel": 2, "tfactor": 3}

    def profile(self, on=True, kernels=None):
        """Live HIP-event timing of the kernel families named in kernels
        (keys of KERNELS; None = all of them).  Enabling resets the records."""
        v = 0
        if on:
            v = 1 if kernels is None else sum(1 << (self.KERNELS[k] + 1) for k in kernels)
        _check(lib().rnpx_profile_enable(self.h, v), "rnpx_profile_enable")

    def profile_read(self, kernel):
        n, ms, fl, by = ctypes.c_long(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().rnpx_profile_read(self.h, self.KERNELS[kernel], ctypes.byref(n), ctypes.byref(ms),
                                       ctypes.byref(fl), ctypes.byref(by)), "rnpx_profile_read")
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}

    def synchronize(self):
        _check(lib().rnpx_synchronize(self.h), "rnpx_synchronize")

    def workspace_bytes(self):
        return lib().rnpx_workspace_bytes(self.h)

    def geqrf(self, a, tau):
        m, n = a.shape
        pre = _tpre(a)
        _check(getattr(lib(), f"rnpx_{pre}geqrf_dev")(self.h, m, n, _dptr(a), _ld(a), _dptr(tau)),
               "geqrf_dev")

    def gelqf(self, a, tau):
        m, n = a.shape
        pre = _tpre(a)
        _check(getattr(lib(), f"rnpx_{pre}gelqf_dev")(self.h, m, n, _dptr(a), _ld(a), _dptr(tau)),
               "gelqf_dev")

    def geqlf(self, a, tau):
        m, n = a.shape
        pre = _tpre(a)
        _check(getattr(lib(), f"rnpx_{pre}geqlf_dev")(self.h, m, n, _dptr(a), _ld(a), _dptr(tau)),
               "geqlf_dev")

    def gerqf(self, a, tau):
        m, n = a.shape
        pre = _tpre(a)
        _check(getattr(lib(), f"rnpx_{pre}gerqf_dev")(self.h, m, n, _dptr(a), _ld(a), _dptr(tau)),
               "gerqf_dev")

    def ormqr(self, side, trans, a, tau, c, k=None):
        m, n = c.shape
        k = tau.shape[0] if k is None else k
        fn = "rnpx_dormqr_dev" if _tpre(c) == "d" else "rnpx_zunmqr_dev"
        _check(getattr(lib(), fn)(self.h, side.encode(), trans.encode(), m, n, k, _dptr(a), _ld(a),
                                  _dptr(tau), _dptr(c), _ld(c)), "ormqr_dev")

    def orgqr(self, a, tau, n=None, k=None):
        m = a.shape[0]
        n = a.shape[1] if n is None else n
        k = min(tau.shape[0], n) if k is None else k
        fn = "rnpx_dorgqr_dev" if _tpre(a) == "d" else "rnpx_zungqr_dev"
        _check(getattr(lib(), fn)(self.h, m, n, k, _dptr(a), _ld(a), _dptr(tau)), "orgqr_dev")

    def gels(self, a, tau, b):
        """min ||A X - B|| in place: A -> factor, B -> Q^H B with X on top."""
        m, n = a.shape
        nrhs = b.shape[1] if b.dim() == 2 else 1
        fn = "rnpx_dgels_dev" if _tpre(a) == "d" else "rnpx_zgels_dev"
        _check(getattr(lib(), fn)(self.h, m, n, nrhs, _dptr(a), _ld(a), _dptr(tau), _dptr(b), _ld(b)), "gels_dev")

    def gels_mixed(self, a, b, x, max_iter=30, tol=1e-14, allow_notconv=False):
        """Mixed-precision LS with FP64 refinement; returns (status, iters, berr)."""
        m, n = a.shape
        nrhs = b.shape[1] if b.dim() == 2 else 1
        it, be = ctypes.c_int(), ctypes.c_double()
        st = lib().rnpx_dgels_mixed_dev(self.h, m, n, nrhs, _dptr(a), _ld(a), _dptr(b), _ld(b), _dptr(x), _ld(x),
                                        max_iter, tol, ctypes.byref(it), ctypes.byref(be))
        if st != RNPX_OK and not (allow_notconv and st == RNPX_ERR_NOTCONV):
            raise RnpxError(st, "gels_mixed_dev")
        return st, it.value, be.value

    def ls_berr(self, a, b, x):
        m, n = a.shape
        nrhs = b.shape[1] if b.dim() == 2 else 1
        be = ctypes.c_double()
        _check(lib().rnpx_dls_berr_dev(self.h, m, n, nrhs, _dptr(a), _ld(a), _dptr(b), _ld(b), _dptr(x), _ld(x),
                                       ctypes.byref(be)), "ls_berr_dev")
        return be.value

    def geqrf_batched(self, desc, A, tau, B=None):
        """Batched QR (and LS where desc['nrhs'] > 0) over packed device buffers."""
        desc = np.ascontiguousarray(desc, dtype=QR_DESC)
        fn = "rnpx_dgeqrf_batched" if _tpre(A) == "d" else "rnpx_zgeqrf_batched"
        _check(getattr(lib(), fn)(self.h, len(desc), desc.ctypes.data_as(_p), _dptr(A), _dptr(tau),
                                  None if B is None else _dptr(B)), "geqrf_batched")

    def geqrf_batched_noalloc(self, desc, A, tau, B=None):
        """rnpx_dgeqrf_batched_noalloc: the batched QR / LS with no device
        allocation (scratch in a static device array, at most 16384
        problems); returns the status code."""
        desc = np.ascontiguousarray(desc, dtype=QR_DESC)
        return lib().rnpx_dgeqrf_batched_noalloc(self.h, len(desc), desc.ctypes.data_as(_p), _dptr(A), _dptr(tau),
                                                 None if B is None else _dptr(B))

    def batched_plan(self, desc, A, tau, B=None):
        desc = np.ascontiguousarray(desc, dtype=QR_DESC)
        plan = ctypes.c_void_p()
        fn = "rnpx_dgeqrf_batched_plan" if _tpre(A) == "d" else "rnpx_zgeqrf_batched_plan"
        _check(getattr(lib(), fn)(self.h, len(desc), desc.ctypes.data_as(_p), _dptr(A), _dptr(tau),
                                  None if B is None else _dptr(B), ctypes.byref(plan)), "batched_plan")
        return BatchedPlan(plan)

    def larft(self, v, tau, t):
        n, k = v.shape
        fn = "rnpx_dlarft_dev" if _tpre(v) == "d" else "rnpx_zlarft_dev"
        _check(getattr(lib(), fn)(self.h, n, k, _dptr(v), _ld(v), _dptr(tau), _dptr(t), _ld(t)), "larft_dev")

    def larfb(self, side, trans, v, t, c):
        m, n = c.shape
        k = v.shape[1]
        fn = "rnpx_dlarfb_dev" if _tpre(c) == "d" else "rnpx_zlarfb_dev"
        _check(getattr(lib(), fn)(self.h, side.encode(), trans.encode(), m, n, k, _dptr(v), _ld(v), _dptr(t),
                                  _ld(t), _dptr(c), _ld(c)), "larfb_dev")

    def gemm(self, transa, transb, alpha, a, b, beta, c):
        m, n = c.shape
        k = a.shape[1] if transa == "N" else a.shape[0]
        if _tpre(c) == "d":
            _check(lib().rnpx_dgemm_dev(self.h, transa.encode(), transb.encode(), m, n, k, float(alpha), _dptr(a),
                                        _ld(a), _dptr(b), _ld(b), float(beta), _dptr(c), _ld(c)), "dgemm_dev")
        else:
            al, be = complex(alpha), complex(beta)
            _check(lib().rnpx_zgemm_dev(self.h, transa.encode(), transb.encode(), m, n, k, ZComplex(al.real, al.imag),
                                        _dptr(a), _ld(a), _dptr(b), _ld(b), ZComplex(be.real, be.imag), _dptr(c),
                                        _ld(c)), "zgemm_dev")

    def larnv(self, dist, iseed, x):
        seed = (_i * 4)(*[int(s) for s in iseed])
        fn = "rnpx_dlarnv_dev" if _tpre(x) == "d" else "rnpx_zlarnv_dev"
        _check(getattr(lib(), fn)(self.h, dist, seed, x.numel(), _dptr(x)), "larnv_dev")


class BatchedPlan:
    """A captured batched launch (hipGraph) bound to its device buffers."""

    def __init__(self, p):
        self.p = p

    def run(self):
        _check(lib().rnpx_batched_plan_run(self.p), "batched_plan_run")

    def close(self):
        if self.p:
            lib().rnpx_batched_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PATHS = {0: "fp64", 1: "mixed", 2: "mixed->fp64", 3: "batched"}


class Server:
    """Least-squares request server (rnpx_server_*): submit host arrays, poll
    for X.  mode 0 = FP64, 1 = mixed precision with FP64 fallback."""

    def __init__(self, hbm_budget=0, workers=4, mode=1, max_iter=30, tol=1e-13):
        cfg = ServerConfig(hbm_budget, workers, mode, max_iter, tol)
        p = ctypes.c_void_p()
        _check(lib().rnpx_server_create(ctypes.byref(cfg), ctypes.byref(p)), "rnpx_server_create")
        self.p = p
        self.shapes = {}

    def submit(self, a, b):
        a = np.asfortranarray(a, dtype=np.float64)
        b2 = np.asfortranarray(b.reshape(b.shape[0], -1), dtype=np.float64)
        m, n = a.shape
        rid = ctypes.c_longlong()
        st = lib().rnpx_server_submit(self.p, m, n, b2.shape[1], _ptr(a), m, _ptr(b2), m, ctypes.byref(rid))
        self.shapes[rid.value] = (n, b2.shape[1])
        return st, rid.value

    def poll(self, rid, wait=True):
        n, q = self.shapes[rid]
        x = np.zeros((n, q), dtype=np.float64, order="F")
        info = RequestInfo()
        st = lib().rnpx_server_poll(self.p, rid, 1 if wait else 0, _ptr(x), n, ctypes.byref(info))
        if st == 1:
            return None, None
        _check(st, "rnpx_server_poll")
        return x, info

    def release(self, rid):
        lib().rnpx_server_release(self.p, rid)
        self.shapes.pop(rid, None)

    def stats(self):
        s = ServerStats()
        _check(lib().rnpx_server_stats_get(self.p, ctypes.byref(s)), "rnpx_server_stats_get")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.p:
            lib().rnpx_server_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
