"""Thin numpy wrappers over the C ABI for the GPU tests (ctypes only)."""
import ctypes

import numpy as np

from oracle_lib import abi

_I = abi.c_int_p


def _boxes(boxes):
    b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1))
    return b, b.ctypes.data_as(_I), len(b) // 4


def struct5(lib, aP, aE, aW, aN, aS, b, boxes, maxITR=30, tol=1e-10, mode=0, x=None):
    ni2, ld = aP.shape
    x = np.zeros_like(aP) if x is None else x
    bk, bp, nb = _boxes(boxes)
    st = abi.Status()
    rc = lib.cathode_struct5_solve(*[abi.dptr(np.ascontiguousarray(a)) for a in (aP, aE, aW, aN, aS, b)],
                                   bp, nb, ni2, ld, ld, maxITR, tol, mode, abi.dptr(x), ctypes.byref(st))
    abi.check(rc, lib)
    return x, st


def sor(lib, aP, aE, aW, aN, aS, b, boxes, alpha, maxITR, tol, x):
    ni2, ld = aP.shape
    bk, bp, nb = _boxes(boxes)
    st = abi.Status()
    rc = lib.cathode_sor_solve(*[abi.dptr(np.ascontiguousarray(a)) for a in (aP, aE, aW, aN, aS, b)],
                               bp, nb, ni2, ld, ld, alpha, maxITR, tol, abi.dptr(x), ctypes.byref(st))
    abi.check(rc, lib)
    return x, st


def helm5(lib, aP, aE, aW, aN, aS, b, boxes, maxITR=100, tol=1e-10, mode=0, x=None):
    """complex128 arrays (ni2, nj2)"""
    ni2, ld = aP.shape
    x = np.zeros_like(aP) if x is None else x
    bk, bp, nb = _boxes(boxes)
    st = abi.Status()
    v = [np.ascontiguousarray(a, dtype=np.complex128).view(np.float64) for a in (aP, aE, aW, aN, aS, b)]
    rc = lib.cathode_helm5_solve(*[abi.dptr(a) for a in v], bp, nb, ni2, ld, ld, maxITR, tol, mode,
                                 abi.dptr(x.view(np.float64)), ctypes.byref(st))
    abi.check(rc, lib)
    return x, st


def etheta(lib, p, ne, J):
    """E (complex, with ghosts), Pabs, P_total, status"""
    E = np.zeros(ne.shape, dtype=np.complex128)
    Pabs = np.zeros(ne.shape)
    Pt = ctypes.c_double()
    st = abi.Status()
    rc = lib.cathode_etheta_solve(ctypes.byref(p), abi.dptr(np.ascontiguousarray(ne)),
                                  abi.dptr(np.ascontiguousarray(J).view(np.float64)),
                                  abi.dptr(E.view(np.float64)), abi.dptr(Pabs), ctypes.byref(Pt),
                                  ctypes.byref(st))
    abi.check(rc, lib)
    return E, Pabs, Pt.value, st


def lu9(lib, coefs, b, boxes, x=None):
    """coefs = (aP, aE, aW, aN, aS, aEN, aES, aWN, aWS)"""
    ni2, ld = b.shape
    x = np.zeros_like(b) if x is None else x
    bk, bp, nb = _boxes(boxes)
    st = abi.Status()
    rc = lib.cathode_lu9_solve(*[abi.dptr(np.ascontiguousarray(a)) for a in coefs], abi.dptr(np.ascontiguousarray(b)),
                               bp, nb, ni2, ld, ld, abi.dptr(x), ctypes.byref(st))
    abi.check(rc, lib)
    return x, st


def ndd_handle(lib, p, st):
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
    abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(abi.fields_struct(st))), lib)
    return h


def ndd_state(lib, h, like):
    out = {k: np.zeros_like(v) for k, v in like.items()}
    abi.check(lib.cathode_ndd_get_state(h, ctypes.byref(abi.fields_struct(out))), lib)
    return out


def ndd_batch(lib, handles, nsteps, workers):
    arr = (ctypes.c_void_p * len(handles))(*[h.value for h in handles])
    reps = (abi.NddReport * len(handles))()
    abi.check(lib.cathode_ndd_batch_step(arr, len(handles), nsteps, workers, reps), lib)
    return list(reps)
