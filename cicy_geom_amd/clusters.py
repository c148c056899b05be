"""Tube clusters, percolation families and junction groups (cicy_cnt_percolation in
libcicy_net.so; Hoshen_Kopelman.cpp:2331-2705, tube part)."""
from __future__ import annotations

import ctypes

import numpy as np

from ._libs import net_lib

FAMILIES = {0: "x", 1: "y", 2: "z", 3: "xy", 4: "xz", 5: "yz", 6: "xyz", -1: None}
_bound = False


def _lib():
    global _bound
    lib = net_lib()
    if not _bound:
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        lib.cicy_cnt_percolation.argtypes = [i64, vp, i64, i64, vp, vp, vp, i64, vp, vp, vp,
                                             ctypes.POINTER(i64), vp, vp, ctypes.POINTER(i64),
                                             vp, vp, ctypes.c_char_p, i64]
        _bound = True
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cnt_percolation(labels, n_labels, boundaries, junctions=None, cnts_inside=None) -> dict:
    """labels, n_labels: labels_cnt from the contact search; boundaries: six lists of tubes
    (boundary_cnt, numbered as the reference's window cut: x faces 4 and 2, y 5 and 3, z 1 and 0);
    junctions: [n, 4] (Pi, tube i, Pj, tube j).  Returns dict(family [n_labels] (-1 when not
    percolated), perc_index [n_labels], n_perc, clusters (tube arrays per label, window order),
    isolated (unlabelled window tubes), junction_cluster [n])."""
    labels = np.ascontiguousarray(labels, np.int32)
    n = len(labels)
    inside = np.ascontiguousarray(np.arange(n) if cnts_inside is None else cnts_inside, np.int64)
    if len(boundaries) != 6:
        raise ValueError("six boundary lists")
    boff = np.zeros(7, np.int64)
    boff[1:] = np.cumsum([len(b) for b in boundaries])
    bcnt = np.ascontiguousarray(np.concatenate([np.asarray(b, np.int64) for b in boundaries])
                                if boff[-1] else np.zeros(1, np.int64), np.int64)
    junc = np.ascontiguousarray(np.zeros((0, 4)) if junctions is None else junctions, np.int64)
    nj = len(junc)
    fam = np.zeros(max(n_labels, 1), np.int32)
    pidx = np.zeros(max(n_labels, 1), np.int32)
    coff = np.zeros(n_labels + 1, np.int64)
    ccnt = np.zeros(max(len(inside), 1), np.int64)
    iso = np.zeros(max(len(inside), 1), np.int64)
    jcl = np.zeros(max(nj, 1), np.int32)
    n_perc, n_iso = ctypes.c_int64(), ctypes.c_int64()
    err = ctypes.create_string_buffer(512)
    ok = _lib().cicy_cnt_percolation(
        n, _p(labels) if n else None, n_labels, len(inside), _p(inside) if len(inside) else None,
        _p(boff), _p(bcnt), nj, _p(junc) if nj else None, _p(fam), _p(pidx),
        ctypes.byref(n_perc), _p(coff), _p(ccnt), ctypes.byref(n_iso), _p(iso), _p(jcl), err, 512)
    if not ok:
        raise RuntimeError(err.value.decode())
    return dict(family=fam[:n_labels], perc_index=pidx[:n_labels], n_perc=n_perc.value,
                clusters=[ccnt[coff[L]:coff[L + 1]] for L in range(n_labels)],
                isolated=iso[:n_iso.value], junction_cluster=jcl[:nj])
