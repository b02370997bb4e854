"""The real RNP reference's blocked QR, built from its own sources -- TEST AND
CPU-BASELINE INFRASTRUCTURE, not the product.

``oracle/_ref/librnpref_qr.so`` (Makefile target ``ref``, from
``ref_qr_driver.cpp`` + the reference's ``src/BLAS_mix.cpp``) links the image's
Fortran-ABI LP64 BLAS, conda MKL (``/opt/conda/lib/libmkl_rt.so.1``), by full
path.  :func:`available` is False when the library was not built (no reference
sources) or MKL cannot be loaded; callers then skip.

Entry points (each forwards to the reference function named):
  factor     RNP::LA::QR::Factor     (RNP/LA/QR.hpp:129-197)
  multq      RNP::LA::QR::MultQ      (RNP/LA/QR.hpp:337-429)
  generateq  RNP::LA::QR::GenerateQ  (RNP/LA/QR.hpp:509-570)
  random_matrix  RNP::MatrixGen::RandomMatrix (RNP/MatrixGen.hpp:120-267)
  lu_factor / lu_invert  RNP::LA::LU::Factor / Invert (RNP/LA/LU.hpp:77-214)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "librnpref_qr.so")
MKL_PATH = "/opt/conda/lib/libmkl_rt.so.1"
_lib = None
_mkl = None
_err = None
_sz = ctypes.c_size_t
_p = ctypes.c_void_p


def _load():
    global _lib, _mkl, _err
    if _lib is not None or _err is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _err = f"{LIB_PATH} not built"
        return None
    if not os.path.exists(MKL_PATH):
        _err = f"{MKL_PATH} not present"
        return None
    # LP64 interface; GNU OpenMP threading (the runtime torch already loads)
    os.environ.setdefault("MKL_INTERFACE_LAYER", "LP64")
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    try:
        _mkl = ctypes.CDLL(MKL_PATH, mode=ctypes.RTLD_GLOBAL)
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        _err = str(e)
        return None
    for pre in ("d", "z"):
        getattr(L, f"refqr_{pre}_factor").argtypes = [_sz, _sz, _p, _sz, _p, _sz]
        getattr(L, f"refqr_{pre}_multq").argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _sz, _sz, _p, _sz, _p,
                                                     _p, _sz]
        getattr(L, f"refqr_{pre}_generateq").argtypes = [_sz, _sz, _sz, _p, _sz, _p]
        getattr(L, f"refqr_{pre}_lu_factor").argtypes = [_sz, _sz, _p, _sz, _p]
        getattr(L, f"refqr_{pre}_lu_factor").restype = ctypes.c_int
        getattr(L, f"refqr_{pre}_lu_invert").argtypes = [_sz, _p, _sz, _p, _sz]
        getattr(L, f"refqr_{pre}_lu_invert").restype = ctypes.c_int
    L.refqr_d_random_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        _sz, _sz, _p, _sz, _p, _p]
    _lib = L
    return _lib


def available():
    return _load() is not None


def why_unavailable():
    _load()
    return _err


def blas_name():
    """The BLAS the reference is linked to, with MKL's own version string."""
    if _load() is None:
        return None
    buf = ctypes.create_string_buffer(256)
    try:
        _mkl.MKL_Get_Version_String(buf, 256)
        v = buf.value.decode(errors="replace").strip()
    except Exception:
        v = "MKL"
    return f"{v} ({MKL_PATH}, LP64, GNU OpenMP threading)"


def set_threads(n):
    """BLAS threads used by the reference's Level-2/3 calls."""
    _load()
    _mkl.MKL_Set_Num_Threads(int(n))


def _pre(dt):
    return "z" if np.dtype(dt).kind == "c" else "d"


def _f(a):
    return np.asfortranarray(a)


def factor(a, lwork_cap=0):
    """QR::Factor on a copy of a; returns (afac, tau)."""
    L = _load()
    a = _f(a).copy(order="F")
    m, n = a.shape
    tau = np.zeros(max(min(m, n), 1), dtype=a.dtype)
    getattr(L, f"refqr_{_pre(a.dtype)}_factor")(m, n, a.ctypes.data, max(m, 1), tau.ctypes.data, lwork_cap)
    return a, tau[:min(m, n)]


def factor_inplace(a, tau):
    """QR::Factor in place on a Fortran-ordered array (timing)."""
    L = _load()
    m, n = a.shape
    getattr(L, f"refqr_{_pre(a.dtype)}_factor")(m, n, a.ctypes.data, max(m, 1), tau.ctypes.data, 0)


def multq(side, trans, afac, tau, c, k=None):
    L = _load()
    afac = _f(afac)
    c = _f(c).copy(order="F")
    m, n = c.shape
    k = len(tau) if k is None else k
    getattr(L, f"refqr_{_pre(c.dtype)}_multq")(side.encode(), trans.encode(), m, n, k, afac.ctypes.data,
                                                max(afac.shape[0], 1), _f(tau).ctypes.data, c.ctypes.data, max(m, 1))
    return c


def generateq(afac, tau, n=None, k=None):
    L = _load()
    m = afac.shape[0]
    n = afac.shape[1] if n is None else n
    k = min(len(tau), n) if k is None else k
    q = _f(afac[:, :n]).copy(order="F")
    getattr(L, f"refqr_{_pre(q.dtype)}_generateq")(m, n, k, q.ctypes.data, max(m, 1), _f(tau).ctypes.data)
    return q


def random_matrix(m, n, cond, mode, dmax, dist, iseed):
    """MatrixGen::RandomMatrix (nonsymmetric, real); returns (a, d, iseed_after)."""
    L = _load()
    a = np.zeros((m, n), dtype=np.float64, order="F")
    d = np.zeros(max(min(m, n), 1))
    s = np.array(iseed, dtype=np.int32)
    L.refqr_d_random_matrix(dist, mode, 0, cond, dmax, m, n, a.ctypes.data, max(m, 1), d.ctypes.data, s.ctypes.data)
    return a, d[:min(m, n)], s


def lu_factor(a):
    L = _load()
    a = _f(a).copy(order="F")
    m, n = a.shape
    piv = np.zeros(max(min(m, n), 1), dtype=np.uint64)
    info = getattr(L, f"refqr_{_pre(a.dtype)}_lu_factor")(m, n, a.ctypes.data, max(m, 1), piv.ctypes.data)
    return a, piv[:min(m, n)], info


def lu_invert(lu, piv, lwork_cap=0):
    """LU::Invert; lwork_cap = n takes the reference's unblocked branch."""
    L = _load()
    a = _f(lu).copy(order="F")
    n = a.shape[0]
    p = np.ascontiguousarray(piv, dtype=np.uint64)
    info = getattr(L, f"refqr_{_pre(a.dtype)}_lu_invert")(n, a.ctypes.data, max(n, 1), p.ctypes.data, lwork_cap)
    return a, info
