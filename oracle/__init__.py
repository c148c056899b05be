"""CPU oracle of the DEA conductance solve -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker.  The product (cicy_geom_amd) never imports it.

* liboracle.so (dea_oracle.c): C restatement of the reference solver in the reference's
  loop and summation order, plus an FP64 truth solver.
* assembly.py: dictionary-based restatement of the reference's std::map assembly, for
  small hand-built networks.
* mpnn.py: float64 numpy restatement of the warm-start network's forward and backward.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_I64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_F64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

ORC_OK, ORC_MAXITER, ORC_BREAKDOWN, ORC_BAD_INPUT = 0, 1, 2, 3


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
    L = ctypes.CDLL(path)
    L.oracle_voltage_vector.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                        ctypes.c_double, _F64]
    L.oracle_spmv_sss.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, _F64, _F64]
    L.oracle_update_residual.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, _F64, _F64]
    L.oracle_pcg_sss.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _I64, _I64,
                                 _F64, _F64, _F64, _F64, _F64, _F64, ctypes.c_int64,
                                 ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_double)]
    L.oracle_cg_truth.restype = ctypes.c_int64
    L.oracle_cg_truth.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, _F64, _F64,
                                  ctypes.c_double, ctypes.c_int64,
                                  ctypes.POINTER(ctypes.c_double)]
    L.oracle_boundary_current.restype = ctypes.c_double
    L.oracle_boundary_current.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64]
    L.oracle_parallel_resistance.restype = ctypes.c_double
    L.oracle_parallel_resistance.argtypes = [ctypes.c_int64, _F64, ctypes.c_double]
    L.oracle_resistance_cnt.restype = ctypes.c_double
    L.oracle_resistance_cnt.argtypes = [_F64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                        ctypes.c_double]
    L.oracle_tunnel_constants.argtypes = [ctypes.c_double] * 4 + [
        ctypes.POINTER(ctypes.c_double)] * 2
    L.oracle_junction_resistance.restype = ctypes.c_double
    L.oracle_junction_resistance.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double]
    _LIB = L
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def voltage_vector(reserved, nodes, r_flag, applied_voltage):
    vef = np.zeros(max(reserved, 1))
    ok = lib().oracle_voltage_vector(reserved, nodes, r_flag, applied_voltage, vef)
    if not ok:
        raise ValueError("invalid R_flag")
    return vef[:reserved]


def spmv_sss(row_ptr, col_ind, values, diag, v):
    n = len(diag)
    out = np.empty(n)
    lib().oracle_spmv_sss(n, _c(row_ptr, np.int64), _c(col_ind, np.int64),
                          _c(values, np.float64), _c(diag, np.float64), _c(v, np.float64),
                          out)
    return out


def update_residual(row_ptr, col_ind, values, diag, x0, r):
    r = _c(r, np.float64).copy()
    lib().oracle_update_residual(len(diag), _c(row_ptr, np.int64), _c(col_ind, np.int64),
                                 _c(values, np.float64), _c(diag, np.float64),
                                 _c(x0, np.float64), r)
    return r


def pcg_sss(nodes, reserved, tol, row_ptr, col_ind, values, diag, r, vef, x,
            max_iter=0):
    """Faithful Solve_DEA_equations_CG_SSS.  Returns dict(voltages, x, r, iters, status,
    rel_res)."""
    r = _c(r, np.float64).copy()
    x = _c(x, np.float64).copy()
    volt = np.empty(nodes)
    it, st, rel = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
    vef = _c(vef, np.float64) if len(vef) else np.zeros(1)
    lib().oracle_pcg_sss(nodes, reserved, tol, _c(row_ptr, np.int64), _c(col_ind, np.int64),
                         _c(values, np.float64), _c(diag, np.float64), r, vef, x, volt,
                         max_iter, ctypes.byref(it), ctypes.byref(st), ctypes.byref(rel))
    return dict(voltages=volt, x=x, r=r, iters=it.value, status=st.value, rel_res=rel.value)


def cg_truth(row_ptr, col_ind, values, diag, b, x0=None, tol=1e-14, max_iter=10_000_000):
    """FP64 truth solve of A x = b to ||b-Ax||/||b|| <= tol."""
    n = len(diag)
    x = np.zeros(n) if x0 is None else _c(x0, np.float64).copy()
    rel = ctypes.c_double()
    k = lib().oracle_cg_truth(n, _c(row_ptr, np.int64), _c(col_ind, np.int64),
                              _c(values, np.float64), _c(diag, np.float64),
                              _c(b, np.float64), x, tol, max_iter, ctypes.byref(rel))
    return x, int(k), rel.value


def boundary_current(n_bnd, n_int, re, voltages):
    return lib().oracle_boundary_current(len(n_bnd), _c(n_bnd, np.int64), _c(n_int, np.int64),
                                         _c(re, np.float64), _c(voltages, np.float64))


def parallel_resistance(rk, r_matrix=float("inf")):
    rk = _c(rk, np.float64)
    return lib().oracle_parallel_resistance(len(rk), rk if len(rk) else np.zeros(1),
                                            r_matrix)


def resistance_cnt(xyz, p1, p2, radius, rho):
    return lib().oracle_resistance_cnt(_c(xyz, np.float64).reshape(-1), p1, p2, radius, rho)


def tunnel_constants(h=6.62606957, e=1.60217662, m=9.10938291, lam=2.0):
    c1, c2 = ctypes.c_double(), ctypes.c_double()
    lib().oracle_tunnel_constants(h, e, m, lam, ctypes.byref(c1), ctypes.byref(c2))
    return c1.value, c2.value


def junction_resistance(dist, l1, cnt1, l2, cnt2, c1, c2):
    return lib().oracle_junction_resistance(dist, l1, int(cnt1), l2, int(cnt2), c1, c2)
