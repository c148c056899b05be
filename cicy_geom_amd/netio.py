"""The reference's network files (Read_Network.cpp readers, Printer.cpp writers) through
libcicy_net.so.  fmt "csv" or "dat"; files live in `directory` (the reference uses the
current directory).

    tubes = read_cnts(d, "csv")        # npts, radii, xyz, structure (point ids per tube)
    plates = read_gnps(d, "dat")       # geo [n, 7] (l, t, theta, phi, cx, cy, cz), rot [n, 3, 3]
    sample = read_sample_geometry(d)   # corner [3], lengths [3]
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._libs import net_lib

_FMT = {"csv": 0, "dat": 1}
_bound = False


def _lib():
    global _bound
    lib = net_lib()
    if not _bound:
        c, i64, f64, vp = ctypes.c_char_p, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        lib.cicy_read_sample_geometry.argtypes = [c, vp, c, i64]
        lib.cicy_read_cnts.argtypes = [c, ctypes.c_int, ctypes.POINTER(i64),
                                       ctypes.POINTER(i64), vp, vp, vp, c, i64]
        lib.cicy_read_gnps.argtypes = [c, ctypes.c_int, ctypes.POINTER(i64), vp, vp, c, i64]
        lib.cicy_write_network_files.argtypes = [c, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 vp, i64, vp, vp, vp, i64, vp, vp, vp, c, i64]
        _bound = True
    return lib


def _fmt(fmt):
    if fmt not in _FMT:
        raise ValueError(f"format must be 'csv' or 'dat', not {fmt!r}")
    return _FMT[fmt]


def _call(fn, *args):
    err = ctypes.create_string_buffer(512)
    if not fn(*args, err, 512):
        raise RuntimeError(err.value.decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def read_sample_geometry(directory):
    """(corner [3], lengths [3]) from sample_geom.csv (Read_Network.cpp:153-199)."""
    out = np.zeros(6)
    _call(_lib().cicy_read_sample_geometry, os.fsencode(directory), _p(out))
    return out[:3], out[3:]


def read_cnts(directory, fmt="csv"):
    """Tubes from cnt_struct.* and cnt_coordinates.* (Read_Network.cpp:201-462):
    dict(npts, radii, xyz [n_points, 3], structure: list of point-id arrays per tube)."""
    lib, f, d = _lib(), _fmt(fmt), os.fsencode(directory)
    n, npt = ctypes.c_int64(), ctypes.c_int64()
    _call(lib.cicy_read_cnts, d, f, ctypes.byref(n), ctypes.byref(npt), None, None, None)
    npts = np.zeros(max(n.value, 1), np.int64)
    radii = np.zeros(max(n.value, 1))
    xyz = np.zeros((max(npt.value, 1), 3))
    _call(lib.cicy_read_cnts, d, f, ctypes.byref(n), ctypes.byref(npt), _p(npts), _p(radii),
          _p(xyz))
    npts, radii, xyz = npts[:n.value], radii[:n.value], xyz[:npt.value]
    off = np.concatenate([[0], np.cumsum(npts)])
    return dict(npts=npts, radii=radii, xyz=xyz,
                structure=[np.arange(off[i], off[i + 1]) for i in range(len(npts))])


def read_gnps(directory, fmt="csv"):
    """Plates from gnp_data.* (Read_Network.cpp:464-678): dict(geo [n, 7] = l, t, theta,
    phi, cx, cy, cz; rot [n, 3, 3] = Get_transformation_matrix(theta, phi))."""
    lib, f, d = _lib(), _fmt(fmt), os.fsencode(directory)
    n = ctypes.c_int64(0)
    _call(lib.cicy_read_gnps, d, f, ctypes.byref(n), None, None)
    geo = np.zeros((max(n.value, 1), 7))
    rot = np.zeros((max(n.value, 1), 3, 3))
    _call(lib.cicy_read_gnps, d, f, ctypes.byref(n), _p(geo), _p(rot))
    return dict(geo=geo[:n.value], rot=rot[:n.value])


def write_network(directory, fmt, corner, lengths, npts=(), radii=(), xyz=None, side=(),
                  thick=(), rot=None, centre=None, prec_cnt=6, prec_gnp=6):
    """The files Generate_Network writes for a network (Generate_Network.cpp:6118-6180,
    Printer.cpp:140-380): sample_geom.csv, the tube files when npts is not empty and the plate
    file when side is not empty."""
    sample = np.ascontiguousarray(np.concatenate([corner, lengths]), np.float64)
    npts = np.ascontiguousarray(npts, np.int64)
    radii = np.ascontiguousarray(radii, np.float64)
    xyz = np.ascontiguousarray(np.zeros((0, 3)) if xyz is None else xyz, np.float64)
    if int(npts.sum()) != len(xyz) or len(radii) != len(npts):
        raise ValueError("npts must sum to the number of points, with one radius per tube")
    lt = np.ascontiguousarray(np.stack([np.asarray(side, float), np.asarray(thick, float)], 1)
                              if len(side) else np.zeros((0, 2)), np.float64)
    rot = np.ascontiguousarray(np.zeros((0, 3, 3)) if rot is None else rot, np.float64)
    centre = np.ascontiguousarray(np.zeros((0, 3)) if centre is None else centre, np.float64)
    if not (len(lt) == len(rot) == len(centre)):
        raise ValueError("one side, thickness, rotation and centre per plate")
    _call(_lib().cicy_write_network_files, os.fsencode(directory), _fmt(fmt), prec_cnt,
          prec_gnp, _p(sample), len(npts), _p(npts), _p(radii), _p(xyz), len(lt), _p(lt),
          _p(rot), _p(centre))


def rotation(theta, phi):
    """Generate_Network::Get_transformation_matrix (Generate_Network.cpp:3466-3490), in numpy
    (for tests and for building plate rotations to write)."""
    ct, st, cp, sp = np.cos(theta), np.sin(theta), np.cos(phi), np.sin(phi)
    return np.array([[cp * ct, -sp, cp * st], [sp * ct, cp, sp * st], [-st, 0.0, ct]])
