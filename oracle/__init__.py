"""ctypes front-end of the oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY.  The oracle is the CPU restatement of IFEM's
assembly + Jacobi-PCG path used as the checker by tests/, by
``__graft_entry__.smoke()`` and by bench.py's ``cpu_baseline`` leg.  The
product package ``ifem_amd`` never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class _Function(C.Structure):
    _fields_ = [("type", C.c_int), ("nmodes", C.c_int), ("scale", C.c_double),
                ("mean", C.c_double), ("modes", C.POINTER(C.c_double))]


class _Integrand(C.Structure):
    _fields_ = [("kind", C.c_int), ("E", C.c_double), ("nu", C.c_double),
                ("kappa", _Function), ("source", _Function), ("body", C.c_double * 3)]


class _Patch(C.Structure):
    _fields_ = [("p", C.c_int * 3), ("n", C.c_int * 3),
                ("knots", C.POINTER(C.c_double) * 3), ("coefs", C.POINTER(C.c_double))]


class _Sam(C.Structure):
    _fields_ = [("meqn", C.POINTER(C.c_int)), ("ndof", C.c_int), ("nceq", C.c_int),
                ("mpmceq", C.POINTER(C.c_int)), ("mmceq", C.POINTER(C.c_int)),
                ("ttcc", C.POINTER(C.c_double))]


def _sam(s, keep):
    arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (s.meqn, s.mpmceq, s.mmceq)]
    tt = np.ascontiguousarray(s.ttcc, dtype=np.float64)
    keep.bufs.extend(arrs + [tt])
    ip = C.POINTER(C.c_int)
    return _Sam(arrs[0].ctypes.data_as(ip), len(arrs[0]), len(arrs[1]) - 1,
                arrs[1].ctypes.data_as(ip), arrs[2].ctypes.data_as(ip),
                tt.ctypes.data_as(C.POINTER(C.c_double)))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_gauss.argtypes = [C.c_int, _dp, _dp]
        L.orc_num_gauss.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_num_gauss.restype = C.c_int
        L.orc_find_span.argtypes = [C.c_int, C.c_int, _dp, C.c_double]
        L.orc_find_span.restype = C.c_int
        L.orc_basis_funs.argtypes = [C.c_int, _dp, C.c_int, C.c_double, _dp, _dp]
        L.orc_jacobian.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_jacobian.restype = C.c_double
        L.orc_element_matrices.argtypes = [C.POINTER(_Patch), C.POINTER(_Integrand), _ip,
                                           C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_pattern.argtypes = [C.POINTER(_Patch), C.c_int, _ip, C.c_int, _lp, C.c_void_p]
        L.orc_pattern.restype = C.c_int64
        L.orc_assemble.argtypes = [C.POINTER(_Patch), C.POINTER(_Integrand), _ip, C.c_int, _ip,
                                   C.c_int, _lp, _ip, _dp, _dp, C.c_int]
        L.orc_assemble_range.argtypes = [C.POINTER(_Patch), C.POINTER(_Integrand), _ip, C.c_int,
                                         _ip, C.c_int, _lp, _ip, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.orc_assemble_range_g.argtypes = [C.POINTER(_Patch), C.POINTER(_Integrand), _ip,
                                           C.c_int, _ip, C.c_int, _lp, _ip, _dp, _dp, C.c_int,
                                           C.c_int, C.c_int, C.c_void_p]
        L.orc_expand_solution_g.argtypes = [C.c_int, _ip, _dp, _dp, _dp]
        L.orc_expand_solution_g.restype = None
        L.orc_pattern_sam.argtypes = [C.POINTER(_Patch), C.c_int, C.POINTER(_Sam), C.c_int, _lp,
                                      C.c_void_p]
        L.orc_pattern_sam.restype = C.c_int64
        L.orc_assemble_sam.argtypes = [C.POINTER(_Patch), C.POINTER(_Integrand), _ip, C.c_int,
                                       C.POINTER(_Sam), C.c_int, _lp, _ip, _dp, _dp]
        L.orc_expand_sam.argtypes = [C.POINTER(_Sam), _dp, _dp]
        L.orc_expand_sam.restype = None
        L.orc_assemble_traction.argtypes = [C.POINTER(_Patch), C.c_int, _dp, C.c_int, C.c_int,
                                            _ip, _dp]
        L.orc_spmv.argtypes = [C.c_int, _lp, _ip, _dp, _dp, _dp]
        L.orc_spmv.restype = None
        L.orc_pcg_jacobi.argtypes = [C.c_int, _lp, _ip, _dp, _dp, _dp, C.c_double, C.c_double,
                                     C.c_int, C.POINTER(C.c_double)]
        L.orc_pcg_jacobi_fixed.argtypes = [C.c_int, _lp, _ip, _dp, _dp, _dp, C.c_int]
        L.orc_pcg_jacobi_fixed.restype = None
        L.orc_eval_function.argtypes = [C.POINTER(_Function), _dp]
        L.orc_eval_function.restype = C.c_double
        L.orc_eval_function_grad.argtypes = [C.POINTER(_Function), _dp, _dp]
        L.orc_eval_function_grad.restype = None
        L.orc_norms.argtypes = [C.POINTER(_Patch), C.POINTER(_Integrand), _ip, _dp,
                                C.c_void_p, _dp]
        _LIB = L
    return _LIB


def ref_lib():
    """oracle/_ref/libifemref.so (reference sources), or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libifemref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_gauss.argtypes = [C.c_int, _dp, _dp]
        L.ref_jacobian.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.ref_jacobian.restype = C.c_double
        L.ref_sam_couplings.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip,
                                        C.POINTER(C.c_int), _lp, _ip, C.c_int64]
        L.ref_sam_couplings.restype = C.c_int64
        L.ref_sam_elmeqns.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int, _ip]
        L.ref_sam_couplings_mpc.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int,
                                            _ip, _ip, _dp, _ip, C.POINTER(C.c_int), _lp, _ip,
                                            C.c_int64, _dp, _dp]
        L.ref_sam_couplings_mpc.restype = C.c_int64
        _REF = L
    return _REF


# ---- marshalling -----------------------------------------------------------
class _Keep:
    """Holds numpy buffers alive while a ctypes struct points at them."""

    def __init__(self):
        self.bufs = []

    def ptr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.bufs.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))


def _patch(P, keep: _Keep) -> _Patch:
    s = _Patch()
    for d in range(3):
        s.p[d] = int(P.p[d])
        s.n[d] = int(P.n[d])
        s.knots[d] = keep.ptr(P.knots[d])
    s.coefs = keep.ptr(P.coefs.reshape(-1))
    return s


def _function(f, keep: _Keep) -> _Function:
    s = _Function()
    s.type = int(f.type)
    s.scale = float(f.scale)
    s.mean = float(f.mean)
    s.nmodes = f.nmodes
    s.modes = keep.ptr(f.modes.reshape(-1)) if f.nmodes else C.POINTER(C.c_double)()
    return s


def _integrand(itg, keep: _Keep) -> _Integrand:
    s = _Integrand()
    s.kind = int(itg.kind)
    s.E = float(itg.E)
    s.nu = float(itg.nu)
    s.kappa = _function(itg.kappa, keep)
    s.source = _function(itg.source, keep)
    for i in range(3):
        s.body[i] = float(itg.body[i])
    return s


# ---- API -------------------------------------------------------------------
def gauss(n: int):
    xi = np.zeros(n); w = np.zeros(n)
    if lib().orc_gauss(n, xi, w):
        raise ValueError(n)
    return xi, w


def num_gauss(order: int, nGauss: int, neumann: bool = False) -> int:
    """ASMbase::getNoGaussPt with the spline order (p+1)."""
    return int(lib().orc_num_gauss(int(order), int(nGauss), 1 if neumann else 0))


def basis(p: int, t: np.ndarray, u: float):
    t = np.ascontiguousarray(t, dtype=np.float64)
    n = len(t) - p - 1
    s = lib().orc_find_span(p, n, t, u)
    N = np.zeros(p + 1); dN = np.zeros(p + 1)
    lib().orc_basis_funs(p, t, s, u, N, dN)
    return s, N, dN


def jacobian(X: np.ndarray, dNdu: np.ndarray):
    nen = dNdu.shape[0]
    Jinv = np.zeros(9); dNdX = np.zeros(nen * 3)
    det = lib().orc_jacobian(nen, np.ascontiguousarray(X.reshape(-1)),
                             np.ascontiguousarray(dNdu.reshape(-1)), Jinv, dNdX)
    return det, Jinv.reshape(3, 3), dNdX.reshape(nen, 3)


def element_matrices(P, itg, ng, e):
    keep = _Keep()
    sp = _patch(P, keep); si = _integrand(itg, keep)
    nedof = itg.ncomp * P.nen
    EK = np.zeros(nedof * nedof); EV = np.zeros(nedof)
    rc = lib().orc_element_matrices(C.byref(sp), C.byref(si), np.asarray(ng, dtype=np.int32),
                                    int(e[0]), int(e[1]), int(e[2]), EK, EV)
    if rc:
        raise RuntimeError("orc_element_matrices failed")
    return EK.reshape(nedof, nedof), EV


def pattern(P, ncomp: int, meqn: np.ndarray, neq: int):
    keep = _Keep()
    sp = _patch(P, keep)
    meqn = np.ascontiguousarray(meqn, dtype=np.int32)
    rowptr = np.zeros(neq + 1, dtype=np.int64)
    nnz = lib().orc_pattern(C.byref(sp), ncomp, meqn, neq, rowptr, None)
    if nnz < 0:
        raise RuntimeError("orc_pattern failed")
    colind = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().orc_pattern(C.byref(sp), ncomp, meqn, neq, rowptr, colind.ctypes.data)
    return rowptr, colind[:nnz]


def assemble(P, itg, ng, meqn, neq, rowptr, colind, nthreads: int = 0,
             e3_range: Optional[tuple] = None, gdof=None):
    """K and b over the patch (or elements e3_range of the third direction).
    gdof: prescribed value per DOF, read where meqn <= 0; the constrained
    columns then move to the right-hand side (src/LinAlg/SparseMatrix.C:639-661)."""
    keep = _Keep()
    sp = _patch(P, keep); si = _integrand(itg, keep)
    meqn = np.ascontiguousarray(meqn, dtype=np.int32)
    A = np.zeros(int(rowptr[-1])); b = np.zeros(neq)
    ngv = np.asarray(ng, dtype=np.int32)
    if gdof is not None:
        g = np.ascontiguousarray(gdof, dtype=np.float64)
        assert g.shape == meqn.shape
        lo, hi = e3_range if e3_range is not None else (0, 1 << 30)
        rc = lib().orc_assemble_range_g(C.byref(sp), C.byref(si), ngv, itg.ncomp, meqn, neq,
                                        rowptr, colind, A, b, nthreads, lo, hi, g.ctypes.data)
    elif e3_range is None:
        rc = lib().orc_assemble(C.byref(sp), C.byref(si), ngv, itg.ncomp, meqn, neq,
                                rowptr, colind, A, b, nthreads)
    else:
        rc = lib().orc_assemble_range(C.byref(sp), C.byref(si), ngv, itg.ncomp, meqn, neq,
                                      rowptr, colind, A, b, nthreads, e3_range[0], e3_range[1])
    if rc:
        raise RuntimeError("orc_assemble failed")
    return A, b


def assemble_into(P, itg, ng, meqn, neq, rowptr, colind, A, b, gdof=None):
    """Adds one patch's element matrices and vectors into A, b (a system of
    several patches: meqn holds the global equation numbers of the patch's
    local DOFs, gdof their prescribed values)."""
    keep = _Keep()
    sp = _patch(P, keep); si = _integrand(itg, keep)
    meqn = np.ascontiguousarray(meqn, dtype=np.int32)
    g = None if gdof is None else np.ascontiguousarray(gdof, dtype=np.float64)
    rc = lib().orc_assemble_range_g(C.byref(sp), C.byref(si), np.asarray(ng, dtype=np.int32),
                                    itg.ncomp, meqn, neq, rowptr, colind, A, b, 0, 0, 1 << 30,
                                    None if g is None else g.ctypes.data)
    if rc:
        raise RuntimeError("orc_assemble_range_g failed")


def pattern_sam(P, ncomp, sam):
    """CSR pattern with constraint equations (masters included)."""
    keep = _Keep()
    sp, ss = _patch(P, keep), _sam(sam, keep)
    rowptr = np.zeros(sam.neq + 1, dtype=np.int64)
    nnz = lib().orc_pattern_sam(C.byref(sp), ncomp, C.byref(ss), sam.neq, rowptr, None)
    if nnz < 0:
        raise RuntimeError("orc_pattern_sam failed")
    colind = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().orc_pattern_sam(C.byref(sp), ncomp, C.byref(ss), sam.neq, rowptr, colind.ctypes.data)
    return rowptr, colind[:nnz]


def assemble_sam(P, itg, ng, sam, rowptr, colind, A=None, b=None):
    """K and b with constraint equations (masters included), added into A, b
    (fresh zeros when not given)."""
    keep = _Keep()
    sp, si, ss = _patch(P, keep), _integrand(itg, keep), _sam(sam, keep)
    A = np.zeros(int(rowptr[-1])) if A is None else A
    b = np.zeros(sam.neq) if b is None else b
    if lib().orc_assemble_sam(C.byref(sp), C.byref(si), np.asarray(ng, dtype=np.int32), itg.ncomp,
                              C.byref(ss), sam.neq, rowptr, colind, A, b):
        raise RuntimeError("orc_assemble_sam failed")
    return A, b


def expand_sam(sam, xeq):
    keep = _Keep()
    ss = _sam(sam, keep)
    out = np.zeros(len(sam.meqn))
    lib().orc_expand_sam(C.byref(ss), np.ascontiguousarray(xeq, dtype=np.float64), out)
    return out


def expand_solution(meqn, xeq, gdof=None):
    """Equation vector -> DOF vector: fixed DOFs 0, or their prescribed value
    from gdof (SAM::expandSolution, src/LinAlg/SAM.C:768-810)."""
    meqn = np.ascontiguousarray(meqn, dtype=np.int32)
    g = np.zeros(len(meqn)) if gdof is None else np.ascontiguousarray(gdof, dtype=np.float64)
    out = np.zeros(len(meqn))
    lib().orc_expand_solution_g(len(meqn), meqn, np.ascontiguousarray(xeq, dtype=np.float64),
                                g, out)
    return out


def add_traction(P, face: int, t, ng: int, ncomp: int, meqn, b):
    keep = _Keep()
    sp = _patch(P, keep)
    rc = lib().orc_assemble_traction(C.byref(sp), face, np.asarray(t, dtype=np.float64), ng,
                                     ncomp, np.ascontiguousarray(meqn, dtype=np.int32), b)
    if rc:
        raise RuntimeError("orc_assemble_traction failed")
    return b


def spmv(rowptr, colind, A, x):
    y = np.zeros(len(rowptr) - 1)
    lib().orc_spmv(len(y), rowptr, colind, A, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def pcg(rowptr, colind, A, b, rtol=1e-10, atol=1e-300, maxits=100000):
    x = np.zeros(len(b))
    res = C.c_double(0.0)
    it = lib().orc_pcg_jacobi(len(b), rowptr, colind, A, b, x, rtol, atol, maxits, C.byref(res))
    return x, it, res.value


def pcg_fixed(rowptr, colind, A, b, iters: int):
    x = np.zeros(len(b))
    lib().orc_pcg_jacobi_fixed(len(b), rowptr, colind, A, b, x, iters)
    return x


def eval_function(f, X):
    keep = _Keep()
    s = _function(f, keep)
    return lib().orc_eval_function(C.byref(s), np.asarray(X, dtype=np.float64))


def eval_function_grad(f, X):
    keep = _Keep()
    s = _function(f, keep)
    g = np.zeros(3)
    lib().orc_eval_function_grad(C.byref(s), np.asarray(X, dtype=np.float64), g)
    return g


def norms(P, itg, ng, xdof, exact=None):
    """[||u_h||_L2, ||u_h||_a, ||u||_L2, ||u||_a, ||u-u_h||_L2, ||u-u_h||_a]."""
    keep = _Keep()
    sp, si = _patch(P, keep), _integrand(itg, keep)
    se = _function(exact, keep) if exact is not None else None
    out = np.zeros(6)
    rc = lib().orc_norms(C.byref(sp), C.byref(si), np.asarray(ng, dtype=np.int32),
                         np.ascontiguousarray(xdof, dtype=np.float64),
                         C.cast(C.pointer(se), C.c_void_p) if se is not None else None, out)
    if rc:
        raise RuntimeError("orc_norms failed")
    return out


def system(P, itg, ng, meqn, neq, traction=None, nthreads: int = 0):
    """Pattern + assembly (+ optional traction (face, t, ng)) -> (rowptr, colind, A, b)."""
    rowptr, colind = pattern(P, itg.ncomp, meqn, neq)
    A, b = assemble(P, itg, ng, meqn, neq, rowptr, colind, nthreads)
    if traction is not None:
        face, t, ngf = traction
        add_traction(P, face, t, ngf, itg.ncomp, meqn, b)
    return rowptr, colind, A, b
