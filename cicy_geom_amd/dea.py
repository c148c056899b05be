"""ctypes binding of libcicy_dea.so (include/cicy_dea.h).

Every call goes to the HIP library; there is no host fallback.  If the library or a GPU
is missing the constructor raises.
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import weakref
from dataclasses import dataclass

import numpy as np

from ._libs import dea_lib

CICY_OK, CICY_MAXITER, CICY_BREAKDOWN, CICY_BAD_INPUT, CICY_DEVICE_ERROR = 0, 1, 2, 3, 4
METHOD_PCG_F64, METHOD_MIXED, METHOD_RESIDENT = 0, 1, 2
METHOD_CONDENSED, METHOD_CONDENSED_RESIDENT, METHOD_CONDENSED_BLOCK = 3, 4, 5
METHOD_NAMES = {0: "pcg_f64", 1: "mixed", 2: "resident", 3: "condensed",
                4: "condensed_resident", 5: "condensed_block"}
FORMAT_AUTO, FORMAT_SSS, FORMAT_BAND = 0, 1, 2
REL_R0, REL_B = 0, 1
STATUS_NAMES = {0: "ok", 1: "maxiter", 2: "breakdown", 3: "bad_input", 4: "device_error"}

_I64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_F64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_VP = ctypes.c_void_p

# Every symbol the header declares; tests check that the library exports all of them.
EXPORTED_SYMBOLS = [
    "cicy_create", "cicy_destroy", "cicy_last_error", "cicy_set_rows_per_block",
    "cicy_set_format", "cicy_host_layout_spmv", "cicy_copy_bench",
    "cicy_solve_dea_cg_sss", "cicy_update_residual", "cicy_spmv_sss_f64",
    "cicy_boundary_current_f64", "cicy_sys_create", "cicy_sys_destroy", "cicy_sys_size",
    "cicy_sys_set_rhs", "cicy_sys_solve", "cicy_sys_get", "cicy_sys_spmv_bench",
    "cicy_batch_create", "cicy_batch_destroy", "cicy_batch_size", "cicy_batch_set_rhs",
    "cicy_batch_solve", "cicy_batch_get", "cicy_mpnn_num_params", "cicy_mpnn_create",
    "cicy_mpnn_destroy", "cicy_mpnn_set_params", "cicy_mpnn_get_params", "cicy_mpnn_get_grads",
    "cicy_mpnn_zero_grads", "cicy_graph_create", "cicy_graph_destroy", "cicy_graph_set_target",
    "cicy_mpnn_forward", "cicy_mpnn_backward", "cicy_mpnn_adam_step", "cicy_mpnn_warm_start",
    "cicy_host_blocked_spmv", "cicy_set_resident_rows", "cicy_host_resident_apply",
    "cicy_host_condense", "cicy_host_expand", "cicy_host_core", "cicy_host_core_expand",
    "cicy_sys_core", "cicy_set_dropin_method", "cicy_sys_set_blocks", "cicy_host_core_block_cg",
    "cicy_sys_prepare",
    "cicy_set_dropin_blocks", "cicy_backbone_scan", "cicy_cnt_contacts",
    "cicy_boundary_current_f64_noalloc", "cicy_sys_spmv_bench_cold", "cicy_read_bench",
    "cicy_sys_create_device", "cicy_sys_set_rhs_device", "cicy_set_setup", "cicy_device_alloc",
    "cicy_device_free", "cicy_device_upload", "cicy_device_download",
]
# Every symbol cicy_assembly.h declares (device assembly, same library).
ASSEMBLY_SYMBOLS = ["cicy_dnet_create", "cicy_dnet_destroy", "cicy_dnet_assemble",
                    "cicy_asm_get", "cicy_asm_free"]


class SolveStats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int64), ("status", ctypes.c_int),
                ("refinements", ctypes.c_int), ("rel_res", ctypes.c_double),
                ("true_rel_res", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("fp64_iters", ctypes.c_int64),
                ("kernel_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_} | {
            "status_name": STATUS_NAMES.get(self.status, str(self.status))}


class JunctionRec(ctypes.Structure):
    """cicy_junction_rec: two points, the two particles they lie on, the gap (um)."""
    _fields_ = [("p1", ctypes.c_int64), ("p2", ctypes.c_int64), ("e1", ctypes.c_int64),
                ("e2", ctypes.c_int64), ("dist", ctypes.c_double)]


def _struct_fields(spec):
    out = []
    for name in spec.split():
        if name.startswith("n_"):
            out.append((name, ctypes.c_int64))
        else:
            out.append((name, _VP))
    return out


class NetArrays(ctypes.Structure):
    """cicy_net_arrays: a network in flat host arrays (cicy_assembly.h)."""
    _fields_ = _struct_fields(
        "n_points xyz n_tubes radius elem_ptr elem_pts n_junctions junctions n_gnp_points gxyz "
        "n_gnps gnp_thickness gstruct_ptr gstruct gtri_ptr gtri n_mixed mixed n_gjunctions "
        "gjunctions bpts_ptr bpts bcnt_ptr bcnt gbpts_ptr gbpts gbnd_ptr gbnd n_clusters "
        "cl_tubes_ptr cl_tubes cl_junc_ptr cl_junc cl_gnps_ptr cl_gnps cl_mixed_ptr cl_mixed "
        "cl_gjunc_ptr cl_gjunc") + [("window_lo", ctypes.c_double * 3),
                                     ("window_hi", ctypes.c_double * 3)]


class AsmView(ctypes.Structure):
    """cicy_asm_view: sizes, host vectors and device arrays of one device assembly."""
    _fields_ = [("nodes", ctypes.c_int64), ("reserved", ctypes.c_int64), ("n", ctypes.c_int64),
                ("nnz", ctypes.c_int64), ("has_guess", ctypes.c_int),
                ("direction", ctypes.c_int), ("vef", ctypes.c_double * 8),
                ("d_row_ptr", _VP), ("d_col_ind", _VP), ("d_values", _VP), ("d_diag", _VP),
                ("d_b", _VP), ("d_x0", _VP), ("d_r0", _VP), ("n_points", ctypes.c_int64),
                ("n_gnp_points", ctypes.c_int64), ("d_lmm", _VP), ("d_lmm_gnp", _VP),
                ("gnp_nodes", ctypes.c_int64), ("n_blocks", ctypes.c_int64), ("blocks", _VP),
                ("ne", ctypes.c_int64 * 2), ("d_e_bnd", _VP * 2), ("d_e_int", _VP * 2),
                ("d_e_re", _VP * 2)]


_bound = None


def lib():
    global _bound
    if _bound is not None:
        return _bound
    L = dea_lib()
    P = ctypes.POINTER
    L.cicy_create.argtypes = [P(_VP), ctypes.c_int]
    L.cicy_destroy.argtypes = [_VP]
    L.cicy_last_error.restype = ctypes.c_char_p
    L.cicy_last_error.argtypes = [_VP]
    L.cicy_set_rows_per_block.argtypes = [_VP, ctypes.c_int]
    L.cicy_set_format.argtypes = [_VP, ctypes.c_int]
    L.cicy_set_resident_rows.argtypes = [_VP, ctypes.c_int]
    L.cicy_set_dropin_method.argtypes = [_VP, ctypes.c_int]
    L.cicy_host_resident_apply.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64,
                                           ctypes.c_int, ctypes.c_int, _F64, _F64,
                                           ctypes.c_double, _F64]
    L.cicy_host_condense.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64,
                                     _VP, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), _I64, _I64, _I64, _F64,
                                     _F64, _VP]
    L.cicy_host_expand.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64, _F64,
                                   _F64, _F64]
    L.cicy_host_core.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64, _VP,
                                 ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_int64), _I64, _I64, _I64, _F64, _F64,
                                 _VP]
    L.cicy_host_core_expand.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64,
                                        _F64, _F64, _F64]
    L.cicy_copy_bench.argtypes = [_VP, ctypes.c_int64, ctypes.c_int, P(ctypes.c_double)]
    L.cicy_read_bench.argtypes = [_VP, ctypes.c_int64, ctypes.c_int, P(ctypes.c_double)]
    L.cicy_sys_spmv_bench_cold.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                           P(ctypes.c_double)]
    L.cicy_host_layout_spmv.argtypes = [ctypes.c_int, ctypes.c_int64, _I64, _I64, _F64, _F64,
                                        ctypes.c_int64, ctypes.c_int, _F64, _F64]
    L.cicy_solve_dea_cg_sss.argtypes = [_VP, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                        _I64, _I64, _F64, _F64, ctypes.c_int64, _F64, _F64, _F64,
                                        _F64, ctypes.c_int64, P(SolveStats)]
    L.cicy_update_residual.argtypes = [_VP, ctypes.c_int64, _I64, _I64, _F64, _F64,
                                       ctypes.c_int64, _F64, _F64]
    L.cicy_spmv_sss_f64.argtypes = [_VP, ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64,
                                    _F64, _F64]
    L.cicy_boundary_current_f64.argtypes = [_VP, ctypes.c_int64, _I64, _I64, _F64, _F64,
                                            ctypes.c_int64, P(ctypes.c_double)]
    L.cicy_boundary_current_f64_noalloc.argtypes = [_VP, ctypes.c_int64, _VP, _VP, _VP, _VP,
                                                    ctypes.c_int64, _VP]
    L.cicy_sys_create.argtypes = [_VP, ctypes.c_int64, ctypes.c_int64, _I64, _I64, _F64, _F64,
                                  P(_VP)]
    L.cicy_sys_create_device.argtypes = [_VP, ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP,
                                         _VP, P(_VP)]
    L.cicy_sys_set_rhs_device.argtypes = [_VP, _VP, _VP, _VP]
    L.cicy_set_setup.argtypes = [_VP, ctypes.c_int]
    L.cicy_device_alloc.argtypes = [_VP, ctypes.c_int64, P(_VP)]
    L.cicy_device_free.argtypes = [_VP, _VP]
    L.cicy_device_upload.argtypes = [_VP, _VP, _VP, ctypes.c_int64]
    L.cicy_device_download.argtypes = [_VP, _VP, _VP, ctypes.c_int64]
    L.cicy_sys_destroy.argtypes = [_VP]
    L.cicy_dnet_create.argtypes = [_VP, _VP, P(_VP)]
    L.cicy_dnet_destroy.argtypes = [_VP]
    L.cicy_dnet_assemble.argtypes = [_VP, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _VP,
                                     ctypes.c_int, P(_VP)]
    L.cicy_asm_get.argtypes = [_VP, P(AsmView)]
    L.cicy_asm_free.argtypes = [_VP]
    L.cicy_sys_size.restype = ctypes.c_int64
    L.cicy_sys_size.argtypes = [_VP, ctypes.c_int]
    L.cicy_sys_set_rhs.argtypes = [_VP, _F64, _VP, _VP]
    L.cicy_sys_solve.argtypes = [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_int64,
                                 ctypes.c_int, P(SolveStats)]
    L.cicy_sys_get.argtypes = [_VP, _VP, _VP]
    L.cicy_sys_core.argtypes = [_VP, P(_VP)]
    L.cicy_sys_set_blocks.argtypes = [_VP, ctypes.c_int64, _VP]
    L.cicy_sys_prepare.argtypes = [_VP, ctypes.c_int]
    L.cicy_set_dropin_blocks.argtypes = [_VP, ctypes.c_int64, _VP]
    L.cicy_host_core_block_cg.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64,
                                          ctypes.c_int64, ctypes.c_int64, _VP, _F64,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                          _F64, P(ctypes.c_int64), P(ctypes.c_int64)]
    L.cicy_sys_spmv_bench.argtypes = [_VP, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
    L.cicy_host_blocked_spmv.argtypes = [ctypes.c_int64, _I64, _I64, _F64, _F64, ctypes.c_int64,
                                         ctypes.c_int, _F64, _F64]
    PP = ctypes.POINTER(ctypes.c_void_p)
    L.cicy_batch_create.argtypes = [_VP, ctypes.c_int, _I64, _I64, PP, PP, PP, PP, P(_VP)]
    L.cicy_batch_destroy.argtypes = [_VP]
    L.cicy_batch_size.restype = ctypes.c_int64
    L.cicy_batch_size.argtypes = [_VP, ctypes.c_int]
    L.cicy_batch_set_rhs.argtypes = [_VP, ctypes.c_int, _F64, _VP, _VP]
    L.cicy_batch_solve.argtypes = [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_int64,
                                   ctypes.c_int, P(SolveStats), P(ctypes.c_double)]
    L.cicy_batch_get.argtypes = [_VP, ctypes.c_int, _VP, _VP]
    _F32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.cicy_mpnn_num_params.restype = ctypes.c_int64
    L.cicy_mpnn_num_params.argtypes = []
    L.cicy_mpnn_create.argtypes = [_VP, _VP, P(_VP)]
    L.cicy_mpnn_destroy.argtypes = [_VP]
    L.cicy_mpnn_set_params.argtypes = [_VP, _F32]
    L.cicy_mpnn_get_params.argtypes = [_VP, _F32]
    L.cicy_mpnn_get_grads.argtypes = [_VP, _F32]
    L.cicy_mpnn_zero_grads.argtypes = [_VP]
    L.cicy_graph_create.argtypes = [_VP, ctypes.c_int64, ctypes.c_int64, _I64, _I64, _F64, _F64,
                                    _F64, _F64, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, P(_VP)]
    L.cicy_graph_destroy.argtypes = [_VP]
    L.cicy_graph_set_target.argtypes = [_VP, _F64]
    L.cicy_mpnn_forward.argtypes = [_VP, _VP, _VP, P(ctypes.c_double)]
    L.cicy_mpnn_backward.argtypes = [_VP, _VP]
    L.cicy_mpnn_adam_step.argtypes = [_VP] + [ctypes.c_double] * 5
    L.cicy_mpnn_warm_start.argtypes = [_VP, _VP, _VP, P(ctypes.c_double)]
    L.cicy_backbone_scan.argtypes = [_VP, ctypes.c_int64, _F64, ctypes.c_int64, _I64, _I64,
                                     ctypes.c_int64, _I64, _I64, P(ctypes.c_double), _VP]
    L.cicy_cnt_contacts.argtypes = [_VP, ctypes.c_int64, _F64, ctypes.c_int64, _F64, _I64,
                                    _I64, ctypes.c_int64, _I64, _F64, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int64, _VP, P(ctypes.c_int64),
                                    ctypes.c_int64, P(ctypes.c_int64), _VP, _VP, _I64, _I64,
                                    _F64]
    _bound = L
    return L


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a if a.size else np.zeros(1, np.int64)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if a.size else np.zeros(1, np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class DEAError(RuntimeError):
    pass


_LIVE_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    """Contexts still open at exit are closed (with their systems, batches and networks)
    while the HIP runtime is certainly up, before the interpreter's teardown; finalizers
    that run later make no native call (see _finalizing)."""
    for ctx in list(_LIVE_CONTEXTS):
        ctx.close()


def _finalizing() -> bool:
    return sys.is_finalizing()


class Context:
    """One device context (stream, events); not re-entrant."""

    def __init__(self, device: int = 0, rows_per_block: int = 0, fmt: int = FORMAT_AUTO):
        L = lib()
        h = _VP()
        if not L.cicy_create(ctypes.byref(h), device):
            raise DEAError(f"cicy_create failed on device {device} (no GPU?)")
        self._h = h
        self._L = L
        self._children = weakref.WeakSet()   # systems, batches, graphs, networks
        _LIVE_CONTEXTS.add(self)
        if rows_per_block:
            self.set_rows_per_block(rows_per_block)
        if fmt:
            self.set_format(fmt)

    def _adopt(self, child):
        """Objects holding device memory of this context are closed before it."""
        self._children.add(child)

    def copy_bench(self, nbytes: int, reps: int = 20) -> float:
        """Average ms of a device copy of nbytes (reads nbytes, writes nbytes)."""
        ms = ctypes.c_double()
        self._check(self._L.cicy_copy_bench(self._h, nbytes, reps, ctypes.byref(ms)),
                    "cicy_copy_bench")
        return ms.value

    def read_bench(self, nbytes: int, reps: int = 20) -> float:
        """Average ms of a read-mostly stream over nbytes."""
        ms = ctypes.c_double()
        self._check(self._L.cicy_read_bench(self._h, nbytes, reps, ctypes.byref(ms)),
                    "cicy_read_bench")
        return ms.value

    def set_format(self, fmt: int):
        self._check(self._L.cicy_set_format(self._h, fmt), "cicy_set_format")

    def set_dropin_method(self, method: int):
        """Method behind solve_dea_cg_sss (default METHOD_PCG_F64, the reference's)."""
        self._check(self._L.cicy_set_dropin_method(self._h, method), "cicy_set_dropin_method")

    def set_dropin_blocks(self, ranges):
        """Row groups [(start, end), ...] (e.g. each nanoplatelet's nodes) for the next
        solve_dea_cg_sss with METHOD_CONDENSED_BLOCK."""
        r = np.ascontiguousarray(np.asarray(ranges, np.int64).reshape(-1, 2))
        self._check(self._L.cicy_set_dropin_blocks(self._h, len(r),
                                                   r.ctypes.data if len(r) else None),
                    "cicy_set_dropin_blocks")

    def set_resident_rows(self, rows_per_thread: int):
        """Rows per thread of the resident layout of systems created afterwards (0: auto)."""
        self._check(self._L.cicy_set_resident_rows(self._h, rows_per_thread),
                    "cicy_set_resident_rows")

    def close(self):
        if getattr(self, "_h", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            self._L.cicy_destroy(self._h)
            self._h = None

    def __del__(self):
        if not _finalizing():
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def last_error(self) -> str:
        return (self._L.cicy_last_error(self._h) or b"").decode()

    def _check(self, ok, what):
        if not ok:
            raise DEAError(f"{what}: {self.last_error}")

    def set_rows_per_block(self, rb: int):
        self._check(self._L.cicy_set_rows_per_block(self._h, rb), "cicy_set_rows_per_block")

    def solve_dea_cg_sss(self, nodes, reserved_nodes, tolerance, col_ind, row_ptr, values,
                         diagonal, R, VEF, voltages_sol, max_iter=0):
        """Drop-in for Solve_DEA_equations_CG_SSS (Direct_Electrifying.cpp:1538-1659).

        Returns (ok, voltages, voltages_sol, R, stats) where R is the final recursive
        residual and voltages = [VEF | x]."""
        n = nodes - reserved_nodes
        R = _f64(np.array(R, dtype=np.float64, copy=True))
        x = _f64(np.array(voltages_sol, dtype=np.float64, copy=True))
        volt = np.zeros(max(nodes, 1))
        st = SolveStats()
        col_ind = _i64(col_ind)
        ok = self._L.cicy_solve_dea_cg_sss(
            self._h, nodes, reserved_nodes, tolerance, col_ind, _i64(row_ptr), _f64(values),
            _f64(diagonal), len(np.asarray(values)), R, _f64(VEF), x, volt, max_iter,
            ctypes.byref(st))
        return bool(ok), volt[:nodes], x[:n], R[:n], st

    def spmv(self, row_ptr, col_ind, values, diagonal, v):
        n = len(diagonal)
        out = np.zeros(max(n, 1))
        self._check(self._L.cicy_spmv_sss_f64(self._h, n, _i64(row_ptr), _i64(col_ind),
                                              _f64(values), _f64(diagonal), len(values),
                                              _f64(v), out), "cicy_spmv_sss_f64")
        return out[:n]

    def update_residual(self, row_ptr, col_ind, values, diagonal, x0, R):
        n = len(diagonal)
        R = _f64(np.array(R, dtype=np.float64, copy=True))
        self._check(self._L.cicy_update_residual(self._h, n, _i64(row_ptr), _i64(col_ind),
                                                 _f64(values), _f64(diagonal), len(values),
                                                 _f64(x0), R), "cicy_update_residual")
        return R[:n]

    def boundary_current(self, n_bnd, n_int, re, voltages) -> float:
        out = ctypes.c_double()
        self._check(self._L.cicy_boundary_current_f64(
            self._h, len(n_bnd), _i64(n_bnd), _i64(n_int), _f64(re), _f64(voltages),
            len(voltages), ctypes.byref(out)), "cicy_boundary_current_f64")
        return out.value

    def boundary_current_noalloc(self, ne, d_n_bnd, d_n_int, d_re, d_voltages, nv, d_current):
        """cicy_boundary_current_f64_noalloc: every argument a device address (int); the
        result is written to d_current on the device."""
        self._check(self._L.cicy_boundary_current_f64_noalloc(
            self._h, ne, d_n_bnd, d_n_int, d_re, d_voltages, nv, d_current),
            "cicy_boundary_current_f64_noalloc")

    def backbone_scan(self, voltages, tube_off, tube_nodes, pair_a, pair_b):
        """Zero current and per-tube backbone runs of a unit-resistor solve on the device
        (Backbone_Network.cpp:64-221, :317-409).  tube_off[t]..tube_off[t + 1] index the
        voltage nodes of tube t's element points in tube_nodes; (pair_a, pair_b) are the
        other elements.  Returns (zero current, int32 array [ntubes, 2] of (idx1, idx2))."""
        off = _i64(tube_off)
        nt = len(np.asarray(tube_off)) - 1
        seg = np.zeros((max(nt, 1), 2), np.int32)
        zero = ctypes.c_double()
        pa = np.asarray(pair_a)
        self._check(self._L.cicy_backbone_scan(
            self._h, len(voltages), _f64(voltages), nt, off, _i64(tube_nodes), len(pa),
            _i64(pa), _i64(pair_b), ctypes.byref(zero), _ptr(seg)), "cicy_backbone_scan")
        return zero.value, seg[:nt]

    def cnt_contacts(self, xyz, radii, structure, window_min, window_len, gs_min, overlap_cnt,
                     overlap_gnp, vdw, tunnel, min_points, cnts_inside=None, as_arrays=False):
        """Tube contacts and cluster labels of the reference's Make_cnt_clusters
        (Hoshen_Kopelman.cpp:203-247) over Contact_grid's subregions (Contact_grid.cpp:18-295).
        structure: list of point-id arrays, one per tube, or a pair (offsets [tubes + 1],
        point ids).  Returns dict(labels, n_labels, junctions [n, 4] (Pi, tube i, Pj, tube j),
        junction_d, elements (list of arrays; with as_arrays, elements_off and elements_pts
        instead), stats)."""
        npts_xyz = np.asarray(xyz).size
        xyz = _f64(np.asarray(xyz, np.float64).reshape(-1))
        n_points = npts_xyz // 3
        if isinstance(structure, tuple):
            off = np.ascontiguousarray(structure[0], np.int64)
            pts = _i64(structure[1])
            n_cnts = len(off) - 1
        else:
            n_cnts = len(structure)
            off = np.zeros(n_cnts + 1, np.int64)
            off[1:] = np.cumsum([len(s) for s in structure])
            pts = _i64(np.concatenate([np.asarray(s, np.int64) for s in structure])
                       if n_cnts else np.zeros(0, np.int64))
        inside = np.arange(n_cnts) if cnts_inside is None else np.asarray(cnts_inside)
        geom = _f64(np.concatenate([window_min, window_len, gs_min, [overlap_cnt, overlap_gnp]]))
        labels = np.zeros(max(n_cnts, 1), np.int32)
        elem_off = np.zeros(n_cnts + 1, np.int64)
        elem_pts = np.zeros(max(int(off[-1]), 1), np.int64)
        stats = np.zeros(8)
        n_labels, n_junc = ctypes.c_int64(), ctypes.c_int64()
        cap = 2 * int(off[-1]) + 1024
        for _ in range(2):
            junc = np.zeros((cap, 4), np.int64)
            jd = np.zeros(cap)
            ok = self._L.cicy_cnt_contacts(
                self._h, n_points, xyz, n_cnts, _f64(radii), off, pts, len(inside),
                _i64(inside), geom, vdw, tunnel, int(min_points), _ptr(labels),
                ctypes.byref(n_labels), cap, ctypes.byref(n_junc), _ptr(junc), _ptr(jd),
                elem_off, elem_pts, stats)
            if ok or n_junc.value <= cap:
                break
            cap = n_junc.value
        self._check(ok, "cicy_cnt_contacts")
        nj = n_junc.value
        out = dict(labels=labels[:n_cnts], n_labels=n_labels.value, junctions=junc[:nj],
                   junction_d=jd[:nj],
                   stats=dict(zip(["subregions", "slots", "pair_tests", "contacts",
                                   "kernel_ms", "host_regions_ms", "device_ms",
                                   "host_labels_ms"], stats.tolist())))
        if as_arrays:
            out.update(elements_off=elem_off, elements_pts=elem_pts[:elem_off[-1]])
        else:
            out["elements"] = [elem_pts[elem_off[i]:elem_off[i + 1]] for i in range(n_cnts)]
        return out

    def system(self, row_ptr, col_ind, values, diagonal) -> "System":
        return System(self, row_ptr, col_ind, values, diagonal)

    def set_setup(self, where: str):
        """Where systems created afterwards are set up: "device" (default) or "host" (the
        host forms, bit-identical; for checking one against the other)."""
        self._check(self._L.cicy_set_setup(self._h, {"device": 0, "host": 1}[where]),
                    "cicy_set_setup")

    def upload(self, a) -> "DeviceArray":
        """A copy of the numpy array `a` in device memory."""
        return DeviceArray(self, a)

    def system_device(self, n: int, nnz: int, d_row_ptr, d_col_ind, d_values,
                      d_diagonal) -> "System":
        """A system from the reference's SSS arrays already in device memory (DeviceArray
        or int addresses; int64 indices): checked and set up on the device."""
        return System(self, None, None, None, None,
                      device=(n, nnz, d_row_ptr, d_col_ind, d_values, d_diagonal))

    def device_network(self, arrays: NetArrays, keep=None) -> "DeviceNetwork":
        """A network copied to the device once (cicy_dnet_create) from its flat arrays;
        `keep` holds whatever owns the host arrays until the copy is done."""
        return DeviceNetwork(self, arrays, keep)

    def batch(self, systems) -> "Batch":
        """systems: sequence of (row_ptr, col_ind, values, diagonal)."""
        return Batch(self, systems)


def _addr(a):
    """Device address of a DeviceArray, an int address, or None."""
    if a is None:
        return None
    return a.ptr if isinstance(a, DeviceArray) else int(a)


class DeviceArray:
    """A numpy array copied into device memory of a context (cicy_device_alloc)."""

    def __init__(self, ctx: Context, a):
        a = np.ascontiguousarray(a)
        self.ctx = ctx
        self.dtype, self.shape, self.nbytes = a.dtype, a.shape, a.nbytes
        p = _VP()
        ctx._check(ctx._L.cicy_device_alloc(ctx._h, a.nbytes, ctypes.byref(p)),
                   "cicy_device_alloc")
        self._p = p
        ctx._adopt(self)
        ctx._check(ctx._L.cicy_device_upload(ctx._h, p, a.ctypes.data_as(_VP), a.nbytes),
                   "cicy_device_upload")

    @property
    def ptr(self) -> int:
        return self._p.value

    def get(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx._L.cicy_device_download(self.ctx._h, out.ctypes.data_as(_VP),
                                                         self._p, self.nbytes),
                        "cicy_device_download")
        return out

    def close(self):
        if getattr(self, "_p", None):
            self.ctx._L.cicy_device_free(self.ctx._h, self._p)
            self._p = None

    def __del__(self):
        if not _finalizing():
            self.close()


class System:
    """A device-resident SSS system (cicy_sys)."""

    def __init__(self, ctx: Context, row_ptr, col_ind, values, diagonal, device=None):
        self.ctx = ctx
        h = _VP()
        if device is not None:
            n, nnz, rp, ci, va, dg = device
            self.n = int(n)
            ctx._check(ctx._L.cicy_sys_create_device(ctx._h, self.n, int(nnz), _addr(rp),
                                                     _addr(ci), _addr(va), _addr(dg),
                                                     ctypes.byref(h)), "cicy_sys_create_device")
        else:
            self.n = len(diagonal)
            ctx._check(ctx._L.cicy_sys_create(ctx._h, self.n, len(values), _i64(row_ptr),
                                              _i64(col_ind), _f64(values), _f64(diagonal),
                                              ctypes.byref(h)), "cicy_sys_create")
        self._h = h
        ctx._adopt(self)

    def close(self):
        if getattr(self, "_h", None):
            self.ctx._L.cicy_sys_destroy(self._h)
            self._h = None

    def __del__(self):
        if not _finalizing():
            self.close()

    def size(self, what: str) -> int:
        return int(self.ctx._L.cicy_sys_size(
            self._h, {"n": 0, "nnz": 1, "remote": 2, "bytes": 3, "format": 4,
                      "off_layout": 5, "resident_rows": 6, "resident_groups": 7,
                      "published": 8, "condensed_rows": 9, "condensed_nnz": 10,
                      "core_rows": 11, "core_nnz": 12, "series_rounds": 13,
                      "core_block_rows": 14, "core_blocks": 15, "core_groups": 16,
                      "core_block_entries": 17}[what]))

    def prepare(self, method):
        """Build the layouts `method` needs now instead of in its first solve."""
        self.ctx._check(self.ctx._L.cicy_sys_prepare(self._h, method), "cicy_sys_prepare")

    def set_blocks(self, ranges):
        """Row groups [(start, end), ...] of free-node rows (e.g. the nodes of each
        nanoplatelet) that METHOD_CONDENSED_BLOCK preconditions with exact block inverses."""
        r = np.ascontiguousarray(np.asarray(ranges, np.int64).reshape(-1, 2))
        self.ctx._check(self.ctx._L.cicy_sys_set_blocks(
            self._h, len(r), r.ctypes.data if len(r) else None), "cicy_sys_set_blocks")

    def set_rhs(self, r0, x0=None, b=None):
        r0 = _f64(r0)
        x0 = None if x0 is None else _f64(x0)
        b = None if b is None else _f64(b)
        self.ctx._check(self.ctx._L.cicy_sys_set_rhs(self._h, r0, _ptr(x0), _ptr(b)),
                        "cicy_sys_set_rhs")
        self._keep = (r0, x0, b)

    def set_rhs_device(self, d_r0, d_x0=None, d_b=None):
        """set_rhs from device arrays (DeviceArray or int addresses)."""
        self.ctx._check(self.ctx._L.cicy_sys_set_rhs_device(self._h, _addr(d_r0), _addr(d_x0),
                                                            _addr(d_b)),
                        "cicy_sys_set_rhs_device")

    def solve(self, method=METHOD_PCG_F64, tol=1e-10, max_iter=0, rel_to=REL_R0) -> SolveStats:
        st = SolveStats()
        self.ctx._check(self.ctx._L.cicy_sys_solve(self._h, method, tol, max_iter, rel_to,
                                                   ctypes.byref(st)), "cicy_sys_solve")
        return st

    def get(self):
        x = np.zeros(max(self.n, 1))
        r = np.zeros(max(self.n, 1))
        self.ctx._check(self.ctx._L.cicy_sys_get(self._h, _ptr(x), _ptr(r)), "cicy_sys_get")
        return x[:self.n], r[:self.n]

    def spmv_bench(self, precision=64, reps=50) -> float:
        ms = ctypes.c_double()
        self.ctx._check(self.ctx._L.cicy_sys_spmv_bench(self._h, precision, reps,
                                                        ctypes.byref(ms)),
                        "cicy_sys_spmv_bench")
        return ms.value

    def spmv_bench_cold(self, precision=64, reps=20, flush_bytes=1 << 30) -> float:
        """Average ms of one SpMV launch with the memory-side cache flushed before it."""
        ms = ctypes.c_double()
        self.ctx._check(self.ctx._L.cicy_sys_spmv_bench_cold(self._h, precision, reps,
                                                             flush_bytes, ctypes.byref(ms)),
                        "cicy_sys_spmv_bench_cold")
        return ms.value

    def core(self):
        """The core system of the condensed method (borrowed: lives as long as this
        system), or None."""
        h = _VP()
        if not self.ctx._L.cicy_sys_core(self._h, ctypes.byref(h)) or not h.value:
            return None
        return _Borrowed(self, h)


class DeviceNetwork:
    """A network in device memory (cicy_dnet): every cluster is assembled from it on the
    device."""

    def __init__(self, ctx: Context, arrays: NetArrays, keep=None):
        self.ctx = ctx
        h = _VP()
        ctx._check(ctx._L.cicy_dnet_create(ctx._h, ctypes.addressof(arrays), ctypes.byref(h)),
                   "cicy_dnet_create")
        self._h = h
        ctx._adopt(self)

    def assemble(self, cluster: int, r_flag: int, family: int, elec=None,
                 direction: int = -1) -> "DeviceAssembly":
        """Cluster `cluster`'s system for (r_flag, family), with the boundary element lists
        of `direction` (0..2; -1 none); elec: network.ElectricParams (default values)."""
        if elec is None:
            from .network import ElectricParams
            elec = ElectricParams()
        ep = elec._c()
        h = _VP()
        self.ctx._check(self.ctx._L.cicy_dnet_assemble(self._h, cluster, r_flag, family,
                                                       ctypes.addressof(ep), direction,
                                                       ctypes.byref(h)), "cicy_dnet_assemble")
        return DeviceAssembly(self.ctx, h)

    def close(self):
        if getattr(self, "_h", None):
            self.ctx._L.cicy_dnet_destroy(self._h)
            self._h = None

    def __del__(self):
        if not _finalizing():
            self.close()


class DeviceAssembly:
    """One cluster's system assembled on the device (cicy_asm): sizes as attributes,
    arrays in device memory until close()."""

    _ARRAYS = {"row_ptr": ("d_row_ptr", np.int64, "n1"), "col_ind": ("d_col_ind", np.int64, "nnz"),
               "values": ("d_values", np.float64, "nnz"), "diag": ("d_diag", np.float64, "n"),
               "b": ("d_b", np.float64, "n"), "x0": ("d_x0", np.float64, "n"),
               "r0": ("d_r0", np.float64, "n"), "lmm": ("d_lmm", np.int64, "n_points"),
               "lmm_gnp": ("d_lmm_gnp", np.int64, "n_gnp_points")}

    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self._h = h
        ctx._adopt(self)
        v = AsmView()
        ctx._check(ctx._L.cicy_asm_get(h, ctypes.byref(v)), "cicy_asm_get")
        self.view = v
        self.nodes, self.reserved, self.n, self.nnz = v.nodes, v.reserved, v.n, v.nnz
        self.has_guess, self.direction = bool(v.has_guess), v.direction
        self.vef = np.array(v.vef[:v.reserved])
        self.gnp_nodes = v.gnp_nodes
        nb = v.n_blocks
        self.blocks = (np.ctypeslib.as_array(ctypes.cast(v.blocks, ctypes.POINTER(ctypes.c_int64)),
                                             (2 * nb,)).reshape(nb, 2).copy()
                       if nb else np.zeros((0, 2), np.int64))

    def _download(self, ptr, dtype, count):
        out = np.empty(count, dtype)
        if count:
            self.ctx._check(self.ctx._L.cicy_device_download(self.ctx._h, out.ctypes.data_as(_VP),
                                                             ptr, out.nbytes),
                            "cicy_device_download")
        return out

    def get(self, name: str) -> np.ndarray:
        """One of row_ptr, col_ind, values, diag, b, x0, r0, lmm, lmm_gnp, copied to host."""
        field, dtype, size = self._ARRAYS[name]
        count = self.n + 1 if size == "n1" else getattr(self.view, size)
        return self._download(getattr(self.view, field), dtype, count)

    def boundary_elements(self):
        """[(n_bnd, n_int, re) for b1, b2] as host arrays (empty without a direction)."""
        v = self.view
        return [(self._download(v.d_e_bnd[w], np.int64, v.ne[w]),
                 self._download(v.d_e_int[w], np.int64, v.ne[w]),
                 self._download(v.d_e_re[w], np.float64, v.ne[w])) for w in range(2)]

    def system(self) -> "System":
        """A solver system on the assembled arrays (no copy through the host), with r0, the
        guess (when the reference uses one) and b set."""
        v = self.view
        s = self.ctx.system_device(v.n, v.nnz, v.d_row_ptr, v.d_col_ind, v.d_values, v.d_diag)
        s.set_rhs_device(v.d_r0, v.d_x0 if v.has_guess else None, v.d_b)
        s._asm = self
        return s

    def close(self):
        if getattr(self, "_h", None):
            self.ctx._L.cicy_asm_free(self._h)
            self._h = None

    def __del__(self):
        if not _finalizing():
            self.close()


class _Borrowed(System):
    """A system handle owned by another system (its core); never destroyed from here."""

    def __init__(self, owner: System, h):
        self.ctx = owner.ctx
        self._owner = owner
        self._h = h
        self.n = int(self.ctx._L.cicy_sys_size(h, 0))

    def close(self):
        self._h = None


class Batch:
    """Many device-resident systems advanced together (cicy_batch)."""

    def __init__(self, ctx: Context, systems):
        self.ctx = ctx
        arrs = [(_i64(rp), _i64(ci), _f64(v), _f64(d)) for rp, ci, v, d in systems]
        self._keep = arrs
        self.sizes = [len(d) for _, _, _, d in systems]
        ns = len(arrs)
        n = np.array(self.sizes, np.int64)
        nnz = np.array([len(v) for _, _, v, _ in systems], np.int64)
        ptrs = [(ctypes.c_void_p * ns)(*[a[k].ctypes.data for a in arrs]) for k in range(4)]
        h = _VP()
        ctx._check(ctx._L.cicy_batch_create(ctx._h, ns, n, nnz, *ptrs, ctypes.byref(h)),
                   "cicy_batch_create")
        self._h = h
        ctx._adopt(self)
        self.nsys = ns

    def close(self):
        if getattr(self, "_h", None):
            self.ctx._L.cicy_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if not _finalizing():
            self.close()

    def size(self, what: str) -> int:
        return int(self.ctx._L.cicy_batch_size(self._h, {
            "n": 0, "nnz": 1, "remote": 2, "bytes": 3, "systems": 4, "blocks": 5,
            "format": 6}[what]))

    def set_rhs(self, sys_id, r0, x0=None, b=None):
        r0 = _f64(r0)
        x0 = None if x0 is None else _f64(x0)
        b = None if b is None else _f64(b)
        self.ctx._check(self.ctx._L.cicy_batch_set_rhs(self._h, sys_id, r0, _ptr(x0), _ptr(b)),
                        "cicy_batch_set_rhs")

    def solve(self, method=METHOD_PCG_F64, tol=1e-10, max_iter=0, rel_to=REL_B):
        st = (SolveStats * self.nsys)()
        ms = ctypes.c_double()
        self.ctx._check(self.ctx._L.cicy_batch_solve(self._h, method, tol, max_iter, rel_to, st,
                                                     ctypes.byref(ms)), "cicy_batch_solve")
        return list(st), ms.value

    def get(self, sys_id):
        n = self.sizes[sys_id]
        x = np.zeros(max(n, 1))
        r = np.zeros(max(n, 1))
        self.ctx._check(self.ctx._L.cicy_batch_get(self._h, sys_id, _ptr(x), _ptr(r)),
                        "cicy_batch_get")
        return x[:n], r[:n]


def host_blocked_spmv(row_ptr, col_ind, values, diagonal, v, rows_per_block=256,
                      fmt=FORMAT_SSS):
    """A device layout applied by host loops (layout check without a GPU)."""
    n = len(diagonal)
    out = np.zeros(max(n, 1))
    ok = lib().cicy_host_layout_spmv(fmt, n, _i64(row_ptr), _i64(col_ind), _f64(values),
                                     _f64(diagonal), len(values), rows_per_block, _f64(v), out)
    if not ok:
        raise DEAError("cicy_host_blocked_spmv rejected the input")
    return out[:n]


def host_resident_apply(row_ptr, col_ind, values, diagonal, r, pold, beta, rows_per_thread,
                        phase):
    """A^ (r + beta pold), A^ = D^-1/2 A D^-1/2, through the resident layout on the host."""
    n = len(diagonal)
    out = np.zeros(max(n, 1))
    ok = lib().cicy_host_resident_apply(n, _i64(row_ptr), _i64(col_ind), _f64(values),
                                        _f64(diagonal), len(values), rows_per_thread, phase,
                                        _f64(r), _f64(pold), float(beta), out)
    if not ok:
        raise DEAError("cicy_host_resident_apply rejected the input")
    return out[:n]


def host_condense(row_ptr, col_ind, values, diagonal, rhs=None):
    """The condensation on the host: (anchors, red_row_ptr, red_col_ind, red_values,
    red_diagonal, red_rhs or None)."""
    n, nnz = len(diagonal), len(values)
    anchors = np.zeros(max(n, 1), np.int64)
    rrp = np.zeros(n + 1, np.int64)
    rcol = np.zeros(max(nnz + n, 1), np.int64)
    rval = np.zeros(max(nnz + n, 1))
    rdiag = np.zeros(max(n, 1))
    rrhs = np.zeros(max(n, 1)) if rhs is not None else None
    m = ctypes.c_int64()
    k = ctypes.c_int64()
    ok = lib().cicy_host_condense(n, _i64(row_ptr), _i64(col_ind), _f64(values), _f64(diagonal),
                                  nnz, _ptr(None if rhs is None else _f64(rhs)), ctypes.byref(m),
                                  ctypes.byref(k), anchors, rrp, rcol, rval, rdiag, _ptr(rrhs))
    if not ok:
        raise DEAError("cicy_host_condense: nothing to condense or invalid input")
    m, k = m.value, k.value
    return (anchors[:m], rrp[:m + 1], rcol[:k], rval[:k], rdiag[:m],
            None if rrhs is None else rrhs[:m])


def host_expand(row_ptr, col_ind, values, diagonal, rhs, x_anchor):
    """Back-substitution of the condensation on the host."""
    n = len(diagonal)
    x = np.zeros(max(n, 1))
    ok = lib().cicy_host_expand(n, _i64(row_ptr), _i64(col_ind), _f64(values), _f64(diagonal),
                                len(values), _f64(rhs), _f64(x_anchor), x)
    if not ok:
        raise DEAError("cicy_host_expand: nothing to condense or invalid input")
    return x[:n]


def host_core_block_cg(row_ptr, col_ind, values, diagonal, rhs, ranges=(), tol=1e-10,
                       ref2=None, max_iter=0):
    """The block-preconditioned core CG (layout + the kernel's recurrences) on the host:
    (x, iterations, blocked rows); raises if it did not converge."""
    n = len(diagonal)
    r = np.ascontiguousarray(np.asarray(ranges, np.int64).reshape(-1, 2))
    x = np.zeros(max(n, 1))
    it, br = ctypes.c_int64(), ctypes.c_int64()
    rhs = _f64(rhs)
    ref2 = float(rhs @ rhs) if ref2 is None else ref2
    ok = lib().cicy_host_core_block_cg(n, _i64(row_ptr), _i64(col_ind), _f64(values),
                                       _f64(diagonal), len(values), len(r),
                                       r.ctypes.data if len(r) else None, rhs, tol, ref2,
                                       max_iter, x, ctypes.byref(it), ctypes.byref(br))
    if not ok:
        raise DEAError("cicy_host_core_block_cg did not converge")
    return x[:n], it.value, br.value


def host_core(row_ptr, col_ind, values, diagonal, rhs=None):
    """Both condensation stages on the host: (core_rows, core_row_ptr, core_col_ind,
    core_values, core_diagonal, core_rhs or None, rounds)."""
    n, nnz = len(diagonal), len(values)
    rows = np.zeros(max(n, 1), np.int64)
    crp = np.zeros(n + 1, np.int64)
    ccol = np.zeros(max(nnz + n, 1), np.int64)
    cval = np.zeros(max(nnz + n, 1))
    cdiag = np.zeros(max(n, 1))
    crhs = np.zeros(max(n, 1)) if rhs is not None else None
    m, k, rounds = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ok = lib().cicy_host_core(n, _i64(row_ptr), _i64(col_ind), _f64(values), _f64(diagonal), nnz,
                              _ptr(None if rhs is None else _f64(rhs)), ctypes.byref(m),
                              ctypes.byref(k), ctypes.byref(rounds), rows, crp, ccol, cval, cdiag,
                              _ptr(crhs))
    if not ok:
        raise DEAError("cicy_host_core: nothing to condense or invalid input")
    m, k = m.value, k.value
    return (rows[:m], crp[:m + 1], ccol[:k], cval[:k], cdiag[:m],
            None if crhs is None else crhs[:m], rounds.value)


def host_core_expand(row_ptr, col_ind, values, diagonal, rhs, x_core):
    """The full solution from the core solution (both stages' back-substitution)."""
    n = len(diagonal)
    x = np.zeros(max(n, 1))
    ok = lib().cicy_host_core_expand(n, _i64(row_ptr), _i64(col_ind), _f64(values),
                                     _f64(diagonal), len(values), _f64(rhs),
                                     _f64(np.asarray(x_core, float).reshape(-1) if len(x_core)
                                          else np.zeros(1)), x)
    if not ok:
        raise DEAError("cicy_host_core_expand: nothing to condense or invalid input")
    return x[:n]
