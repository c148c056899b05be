"""Synthetic CNT/GNP networks and DEA system assembly (Python side of libcicy_net.so).

The generator is this package's own (straight tubes and square nanoplatelets, seeded
mt19937_64); the assembly
follows Direct_Electrifying::Compute_voltage_field (Direct_Electrifying.cpp:11-113) and
its helpers, so the arrays handed to the solver have the reference's layout:

* ``row_ptr``/``col_ind``/``values``: strictly-lower SSS triangle of the free block
  (From_2d_to_1d_vectors, Direct_Electrifying.cpp:1329-1394), int64 indices;
* ``diag``: free-node diagonal; ``b``: -K_EF^T VEF; ``vef``: prescribed voltages;
* ``x0``/``r0``: the ramp warm start and b - A x0 (Direct_Electrifying.cpp:80-95).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from ._libs import net_lib

I64P = ctypes.POINTER(ctypes.c_int64)
F64P = ctypes.POINTER(ctypes.c_double)


class _NetParams(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double), ("n_cnt", ctypes.c_int64),
                ("radius", ctypes.c_double), ("len_min", ctypes.c_double),
                ("len_max", ctypes.c_double), ("seg_len", ctypes.c_double),
                ("d_tunnel", ctypes.c_double), ("d_vdw", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("per_point_nodes", ctypes.c_int),
                ("n_gnp", ctypes.c_int64), ("gnp_side", ctypes.c_double),
                ("gnp_thick", ctypes.c_double), ("min_points", ctypes.c_int)]


class _ElecParams(ctypes.Structure):
    _fields_ = [("applied_voltage", ctypes.c_double), ("resistivity_cnt", ctypes.c_double),
                ("junction_exponential", ctypes.c_int),
                ("junction_resistance", ctypes.c_double),
                ("c1", ctypes.c_double), ("c2", ctypes.c_double),
                ("resistivity_gnp_t", ctypes.c_double),
                ("resistivity_gnp_surf", ctypes.c_double)]


class _SSSView(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int64), ("reserved", ctypes.c_int64),
                ("n", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("row_ptr", I64P), ("col_ind", I64P), ("values", F64P), ("diag", F64P),
                ("b", F64P), ("vef", F64P), ("x0", F64P), ("r0", F64P),
                ("has_guess", ctypes.c_int), ("n_points", ctypes.c_int64),
                ("lmm", I64P), ("node_point", I64P),
                ("ne", ctypes.c_int64 * 2), ("e_bnd", I64P * 2), ("e_int", I64P * 2),
                ("e_re", F64P * 2), ("direction", ctypes.c_int), ("node_xyz", F64P),
                ("gnp_nodes", ctypes.c_int64), ("n_gnp_points", ctypes.c_int64),
                ("lmm_gnp", I64P), ("priv", ctypes.c_void_p)]


_bound = False


def _bind():
    global _bound
    lib = net_lib()
    if _bound:
        return lib
    lib.cicy_net_generate.restype = ctypes.c_void_p
    lib.cicy_net_generate.argtypes = [ctypes.POINTER(_NetParams)]
    lib.cicy_net_free.argtypes = [ctypes.c_void_p]
    lib.cicy_net_count.restype = ctypes.c_int64
    lib.cicy_net_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.cicy_net_xyz.restype = F64P
    lib.cicy_net_xyz.argtypes = [ctypes.c_void_p]
    lib.cicy_net_cnt_ptr.restype = I64P
    lib.cicy_net_cnt_ptr.argtypes = [ctypes.c_void_p]
    lib.cicy_net_cluster.argtypes = [ctypes.c_void_p, ctypes.c_int64, I64P,
                                     ctypes.POINTER(ctypes.c_int)]
    lib.cicy_net_junction.argtypes = [ctypes.c_void_p, ctypes.c_int64, I64P, I64P, I64P,
                                      I64P, F64P]
    lib.cicy_net_boundary_points.restype = ctypes.c_int64
    lib.cicy_net_boundary_points.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(I64P)]
    for nm in ("cicy_net_boundary_tubes", "cicy_net_cluster_tubes",
               "cicy_net_cluster_junctions", "cicy_net_elements", "cicy_net_gnp_structure",
               "cicy_net_gnp_triangulation", "cicy_net_cluster_gnps",
               "cicy_net_cluster_mixed_junctions", "cicy_net_cluster_gnp_junctions"):
        f = getattr(lib, nm)
        f.restype = ctypes.c_int64
        f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(I64P)]
    for nm in ("cicy_net_gnp_boundary_points", "cicy_net_boundary_gnps"):
        f = getattr(lib, nm)
        f.restype = ctypes.c_int64
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(I64P)]
    lib.cicy_net_gnp_xyz.restype = F64P
    lib.cicy_net_gnp_xyz.argtypes = [ctypes.c_void_p]
    lib.cicy_net_gnp.argtypes = [ctypes.c_void_p, ctypes.c_int64, F64P, F64P, F64P, F64P]
    for nm in ("cicy_net_mixed_junction", "cicy_net_gnp_junction"):
        getattr(lib, nm).argtypes = [ctypes.c_void_p, ctypes.c_int64, I64P, I64P, I64P, I64P,
                                     F64P]
    lib.cicy_net_assemble.restype = ctypes.POINTER(_SSSView)
    lib.cicy_net_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(_ElecParams),
                                      ctypes.c_int, ctypes.c_char_p, ctypes.c_int64]
    lib.cicy_sss_free.argtypes = [ctypes.POINTER(_SSSView)]
    lib.cicy_net_arrays_get.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.cicy_net_backbone.argtypes = [ctypes.c_void_p, ctypes.c_int64, F64P, ctypes.c_int64,
                                      I64P, ctypes.c_int64, I64P, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, F64P,
                                      ctypes.c_char_p, ctypes.c_int64]
    lib.cicy_tunnel_constants.argtypes = [ctypes.c_double] * 4 + [F64P, F64P]
    _bound = True
    return lib


def tunnel_constants(h=6.62606957, e=1.60217662, m=9.10938291, lam=2.0):
    """C1, C2 of the exponential junction law (Input_Reader.cpp:1450-1456)."""
    lib = _bind()
    c1, c2 = ctypes.c_double(), ctypes.c_double()
    lib.cicy_tunnel_constants(h, e, m, lam, ctypes.byref(c1), ctypes.byref(c2))
    return c1.value, c2.value


@dataclass
class ElectricParams:
    """Electrical parameters of the reference's Electric_para (Input_Reader.h)."""
    applied_voltage: float = 1.0
    resistivity_cnt: float = 20.0          # Ohm.um (input.dat)
    junction_exponential: bool = True
    junction_resistance: float = 1e7       # used when junction_exponential is False
    lambda_barrier: float = 2.0            # eV
    c1: float = field(default=0.0)
    c2: float = field(default=0.0)
    resistivity_gnp_t: float = 1e-2        # Ohm.um across a plate (input.dat)
    resistivity_gnp_surf: float = 1e-7     # Ohm.um in a plate (input.dat)
    resistivity_matrix: float = 1.03e20    # Ohm.um, the polymer (input.dat)

    def __post_init__(self):
        if self.c1 == 0.0 and self.c2 == 0.0:
            self.c1, self.c2 = tunnel_constants(lam=self.lambda_barrier)

    def _c(self) -> _ElecParams:
        return _ElecParams(self.applied_voltage, self.resistivity_cnt,
                           1 if self.junction_exponential else 0,
                           self.junction_resistance, self.c1, self.c2,
                           self.resistivity_gnp_t, self.resistivity_gnp_surf)


@dataclass
class NetworkParams:
    box: float = 1.0
    n_cnt: int = 300
    radius: float = 0.005
    len_min: float = 0.5
    len_max: float = 1.0
    seg_len: float = 0.02
    d_tunnel: float = 0.0014
    d_vdw: float = 0.00034
    seed: int = 1_000_000
    per_point_nodes: bool = True
    n_gnp: int = 0                # square nanoplatelets
    gnp_side: float = 1.0         # um
    gnp_thick: float = 0.01       # um
    min_points: int = 5           # tube-plate contact compression (input.dat cutoffs)


def config_params(config: int, realisation: int = 0, n_cnt: int | None = None) -> NetworkParams:
    """Network parameters of the workload configurations.

    1: 1 um RVE, 300 tubes;  2: 5 um, 40k tubes;  3: 8 um, 160k tubes + 2k plates;
    4: batch realisations with U[2k,40k] tubes at the density of configuration 2.
    The seed is 1_000_000*config + realisation.
    """
    seed = 1_000_000 * config + realisation
    if config == 1:
        return NetworkParams(box=1.0, n_cnt=300, seed=seed)
    if config == 2:
        return NetworkParams(box=5.0, n_cnt=40_000, seed=seed)
    if config == 3:
        return NetworkParams(box=8.0, n_cnt=160_000, n_gnp=2_000, seed=seed)
    if config in (4, 5):
        if n_cnt is None:
            rng = np.random.default_rng(seed)
            n_cnt = int(rng.integers(2_000, 40_001))
        box = 5.0 * (n_cnt / 40_000.0) ** (1.0 / 3.0)
        return NetworkParams(box=box, n_cnt=n_cnt, seed=seed)
    raise ValueError(f"unknown configuration {config}")


@dataclass
class DEASystem:
    """One assembled DEA system (the inputs of Solve_DEA_equations_CG_SSS)."""
    nodes: int
    reserved: int
    row_ptr: np.ndarray
    col_ind: np.ndarray
    values: np.ndarray
    diag: np.ndarray
    b: np.ndarray
    vef: np.ndarray
    x0: np.ndarray
    r0: np.ndarray
    has_guess: bool
    lmm: np.ndarray
    node_point: np.ndarray
    boundary_elements: list  # [(n_bnd, n_int, re) for b1, b2]
    direction: int
    lmm_gnp: np.ndarray | None = None    # plate point -> global node or -1
    gnp_nodes: int = 0                   # free nodes that are plate points (numbered last)
    node_xyz: np.ndarray | None = None

    @property
    def n(self) -> int:
        return self.nodes - self.reserved

    @property
    def nnz(self) -> int:
        return int(self.col_ind.shape[0])


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Network:
    """A generated network (tubes, junctions, clusters, boundary point lists)."""

    def __init__(self, params: NetworkParams):
        lib = _bind()
        p = _NetParams(params.box, params.n_cnt, params.radius, params.len_min,
                       params.len_max, params.seg_len, params.d_tunnel, params.d_vdw,
                       params.seed, 1 if params.per_point_nodes else 0, params.n_gnp,
                       params.gnp_side, params.gnp_thick, params.min_points)
        self.params = params
        self._h = lib.cicy_net_generate(ctypes.byref(p))
        if not self._h:
            raise ValueError("invalid network parameters")
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cicy_net_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def count(self, what: str) -> int:
        return int(self._lib.cicy_net_count(
            self._h, {"points": 0, "tubes": 1, "junctions": 2, "clusters": 3, "gnps": 4,
                      "gnp_points": 5, "mixed_junctions": 6, "gnp_junctions": 7}[what]))

    @property
    def xyz(self) -> np.ndarray:
        n = self.count("points")
        return _arr(self._lib.cicy_net_xyz(self._h), 3 * n, np.float64).reshape(n, 3)

    @property
    def cnt_ptr(self) -> np.ndarray:
        return _arr(self._lib.cicy_net_cnt_ptr(self._h), self.count("tubes") + 1, np.int64)

    def cluster(self, k: int):
        nt, fam = ctypes.c_int64(), ctypes.c_int()
        if not self._lib.cicy_net_cluster(self._h, k, ctypes.byref(nt), ctypes.byref(fam)):
            raise IndexError(k)
        return nt.value, fam.value

    def junction(self, j: int):
        v = [ctypes.c_int64() for _ in range(4)] + [ctypes.c_double()]
        if not self._lib.cicy_net_junction(self._h, j, *[ctypes.byref(x) for x in v]):
            raise IndexError(j)
        return tuple(x.value for x in v)

    def boundary_points(self, b: int) -> np.ndarray:
        ptr = I64P()
        n = self._lib.cicy_net_boundary_points(self._h, b, ctypes.byref(ptr))
        return _arr(ptr, n, np.int64)

    def _list(self, fn, k) -> np.ndarray:
        ptr = I64P()
        n = fn(self._h, k, ctypes.byref(ptr))
        if n < 0:
            raise IndexError(k)
        return _arr(ptr, n, np.int64)

    def boundary_tubes(self, b: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_boundary_tubes, b)

    def cluster_tubes(self, k: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_cluster_tubes, k)

    def cluster_junctions(self, k: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_cluster_junctions, k)

    def elements(self, t: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_elements, t)

    # ---- nanoplatelets
    @property
    def gnp_xyz(self) -> np.ndarray:
        n = self.count("gnp_points")
        return _arr(self._lib.cicy_net_gnp_xyz(self._h), 3 * n, np.float64).reshape(n, 3)

    def gnp(self, g: int):
        """(centre[3], axes[3,3] with the normal in row 2, side, thickness) of plate g."""
        c = (ctypes.c_double * 3)()
        ax = (ctypes.c_double * 9)()
        l, t = ctypes.c_double(), ctypes.c_double()
        if not self._lib.cicy_net_gnp(self._h, g, c, ax, ctypes.byref(l), ctypes.byref(t)):
            raise IndexError(g)
        return np.array(c[:]), np.array(ax[:]).reshape(3, 3), l.value, t.value

    def gnp_structure(self, g: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_gnp_structure, g)

    def gnp_triangulation(self, g: int) -> np.ndarray:
        ptr = I64P()
        n = self._lib.cicy_net_gnp_triangulation(self._h, g, ctypes.byref(ptr))
        if n < 0:
            raise IndexError(g)
        return _arr(ptr, 2 * n, np.int64).reshape(n, 2)

    def gnp_boundary_points(self, b: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_gnp_boundary_points, b)

    def boundary_gnps(self, b: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_boundary_gnps, b)

    def mixed_junction(self, j: int):
        v = [ctypes.c_int64() for _ in range(4)] + [ctypes.c_double()]
        if not self._lib.cicy_net_mixed_junction(self._h, j, *[ctypes.byref(x) for x in v]):
            raise IndexError(j)
        return tuple(x.value for x in v)

    def gnp_junction(self, j: int):
        v = [ctypes.c_int64() for _ in range(4)] + [ctypes.c_double()]
        if not self._lib.cicy_net_gnp_junction(self._h, j, *[ctypes.byref(x) for x in v]):
            raise IndexError(j)
        return tuple(x.value for x in v)

    def cluster_gnps(self, k: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_cluster_gnps, k)

    def cluster_mixed_junctions(self, k: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_cluster_mixed_junctions, k)

    def cluster_gnp_junctions(self, k: int) -> np.ndarray:
        return self._list(self._lib.cicy_net_cluster_gnp_junctions, k)

    @property
    def radii(self) -> np.ndarray:
        return np.full(self.count("tubes"), self.params.radius)

    def percolating_clusters(self, direction: int) -> list[int]:
        """Clusters whose family includes `direction` (0 x, 1 y, 2 z)."""
        inc = {0: {0, 3, 4, 6}, 1: {1, 3, 5, 6}, 2: {2, 4, 5, 6}}[direction]
        return [k for k in range(self.count("clusters")) if self.cluster(k)[1] in inc]

    def backbone_scan_inputs(self, cluster: int, system: "DEASystem"):
        """The arrays dea.Context.backbone_scan reads for `cluster` and its unit-resistor
        system: (tube_off, tube_nodes) -- the voltage nodes of each cluster tube's element
        points, in cluster order -- and (pair_a, pair_b), the nodes of the other elements the
        zero current is taken over (Backbone_Network.cpp:64-221): tube junctions, plate
        triangulation edges, plate junctions and mixed junctions."""
        lmm = np.asarray(system.lmm, np.int64)
        lg = np.asarray(system.lmm_gnp, np.int64)
        tubes = self.cluster_tubes(cluster)
        els = [self.elements(int(t)) for t in tubes]
        off = np.zeros(len(els) + 1, np.int64)
        off[1:] = np.cumsum([len(e) for e in els])
        nodes = lmm[np.concatenate(els)] if els else np.zeros(0, np.int64)
        pa, pb = [], []
        for j in self.cluster_junctions(cluster):
            p1, p2 = self.junction(int(j))[:2]
            pa.append(lmm[p1])
            pb.append(lmm[p2])
        tri = [self.gnp_triangulation(int(g)) for g in self.cluster_gnps(cluster)]
        for j in self.cluster_gnp_junctions(cluster):
            p1, p2 = self.gnp_junction(int(j))[:2]
            pa.append(lg[p1])
            pb.append(lg[p2])
        for j in self.cluster_mixed_junctions(cluster):
            pc, pg = self.mixed_junction(int(j))[:2]
            pa.append(lmm[pc])
            pb.append(lg[pg])
        pa = np.array(pa, np.int64)
        pb = np.array(pb, np.int64)
        if tri:
            e = np.concatenate(tri)
            pa = np.concatenate([pa, lg[e[:, 0]]])
            pb = np.concatenate([pb, lg[e[:, 1]]])
        return off, nodes, pa, pb

    def backbone(self, cluster: int, system: "DEASystem", voltages, avoid_resistance=False,
                 zero=None, segments=None) -> dict:
        """Backbone of `cluster` (Backbone_Network::Determine_backbone_network,
        Backbone_Network.cpp:10-61) from the voltages [VEF | x] of its unit-resistor system
        `system` (an R_flag = 0 assembly of this cluster).  Edits the cluster as the reference
        does, so the next assemble() gives the backbone's system.  `zero` and `segments`
        (int32 [2 per cluster tube]: first and end backbone element, -1 when none) may come
        from the device scan (dea.Context.backbone_scan)."""
        v = np.ascontiguousarray(voltages, np.float64)
        lmm = np.ascontiguousarray(system.lmm, np.int64)
        lg = np.ascontiguousarray(system.lmm_gnp, np.int64)
        out = np.zeros(5)
        z = None if zero is None else ctypes.c_double(zero)
        seg = None if segments is None else np.ascontiguousarray(segments, np.int32)
        err = ctypes.create_string_buffer(512)
        ok = self._lib.cicy_net_backbone(
            self._h, cluster, v.ctypes.data_as(F64P), len(v), lmm.ctypes.data_as(I64P), len(lmm),
            lg.ctypes.data_as(I64P), len(lg), system.reserved, int(bool(avoid_resistance)),
            None if z is None else ctypes.addressof(z),
            None if seg is None else seg.ctypes.data, out.ctypes.data_as(F64P), err, 512)
        if not ok:
            raise RuntimeError(err.value.decode())
        return dict(zero_current=out[0], cnt_backbone=out[1], cnt_dead=out[2],
                    gnp_backbone=out[3], gnp_dead=out[4])

    def device(self, ctx):
        """This network copied to the device of `ctx` (dea.DeviceNetwork) for device
        assembly; the copy does not follow later changes (backbone trimming)."""
        from .dea import NetArrays
        a = NetArrays()
        if not self._lib.cicy_net_arrays_get(self._h, ctypes.addressof(a)):
            raise RuntimeError("cicy_net_arrays_get failed")
        return ctx.device_network(a, keep=self)

    def assemble(self, cluster: int, r_flag: int, family: int,
                 elec: ElectricParams | None = None, direction: int = -1,
                 with_xyz: bool = False) -> DEASystem:
        elec = elec or ElectricParams()
        ep = elec._c()
        err = ctypes.create_string_buffer(512)
        vp = self._lib.cicy_net_assemble(self._h, cluster, r_flag, family, ctypes.byref(ep),
                                         direction, err, 512)
        if not vp:
            raise RuntimeError(err.value.decode())
        v = vp.contents
        try:
            n, nnz, nodes, rd = v.n, v.nnz, v.nodes, v.reserved
            sys_ = DEASystem(
                nodes=nodes, reserved=rd,
                row_ptr=_arr(v.row_ptr, n + 1, np.int64),
                col_ind=_arr(v.col_ind, nnz, np.int64),
                values=_arr(v.values, nnz, np.float64),
                diag=_arr(v.diag, n, np.float64),
                b=_arr(v.b, n, np.float64),
                vef=_arr(v.vef, rd, np.float64),
                x0=_arr(v.x0, n, np.float64),
                r0=_arr(v.r0, n, np.float64),
                has_guess=bool(v.has_guess),
                lmm=_arr(v.lmm, v.n_points, np.int64),
                node_point=_arr(v.node_point, nodes, np.int64),
                boundary_elements=[(_arr(v.e_bnd[w], v.ne[w], np.int64),
                                    _arr(v.e_int[w], v.ne[w], np.int64),
                                    _arr(v.e_re[w], v.ne[w], np.float64)) for w in range(2)],
                direction=v.direction,
                lmm_gnp=_arr(v.lmm_gnp, v.n_gnp_points, np.int64),
                gnp_nodes=int(v.gnp_nodes))
            if with_xyz:
                sys_.node_xyz = _arr(v.node_xyz, 3 * nodes, np.float64).reshape(nodes, 3)[rd:]
        finally:
            self._lib.cicy_sss_free(vp)
        return sys_


def generate(config: int, realisation: int = 0, **overrides) -> Network:
    p = config_params(config, realisation, overrides.pop("n_cnt", None))
    for k, v in overrides.items():
        setattr(p, k, v)
    return Network(p)


def largest_percolating_system(net: Network, direction: int = 0, r_flag: int = 1,
                               elec: ElectricParams | None = None,
                               with_xyz: bool = False) -> DEASystem:
    """Resistance solve of the largest cluster percolating along `direction`, as
    Electrical_resistance_along_each_percolated_direction sets it up (family :=
    direction, R_flag=1; Electrical_analysis.cpp:249-265)."""
    ks = net.percolating_clusters(direction)
    if not ks:
        raise RuntimeError("no cluster percolates along direction %d" % direction)
    k = max(ks, key=lambda c: net.cluster(c)[0])
    return net.assemble(k, r_flag, direction, elec, direction=direction, with_xyz=with_xyz)


def plate_row_ranges(net: Network, cluster: int, s: DEASystem) -> np.ndarray:
    """[(start, end), ...] free-node rows of each nanoplatelet of the cluster: the
    reference numbers a plate's points consecutively (LM_matrix_for_gnps,
    Direct_Electrifying.cpp:307-349), so each plate is one contiguous range."""
    out = []
    for g in net.cluster_gnps(cluster):
        nodes = s.lmm_gnp[net.gnp_structure(int(g))]
        nodes = nodes[nodes >= s.reserved] - s.reserved
        if len(nodes):
            lo, hi = int(nodes.min()), int(nodes.max()) + 1
            assert hi - lo == len(nodes), "plate nodes are not contiguous"
            out.append((lo, hi))
    return np.array(out, np.int64).reshape(-1, 2)


def wavy_tubes(seed: int, n_tubes: int, box: float, step: float = 0.01, n_min: int = 20,
               n_max: int = 80, r=(0.004, 0.006), kink: float = 0.09):
    """Synthetic wavy tubes for the point-based contact search: each starts at a uniform
    point with an isotropic direction and grows up to U[n_min, n_max) steps of `step` um,
    turning by a Gaussian kick of `kink` per step, and ends at its first point outside
    [0, box]^3 (as a window cut would).  Tubes with fewer than two points are dropped.
    Returns (xyz [points, 3], radii [tubes], structure: point-id ranges per tube, consecutive
    ids), the layout the network file readers produce."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.0, box, (n_tubes, 3))
    d = rng.standard_normal((n_tubes, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    steps = rng.integers(n_min, n_max, n_tubes)
    radii = rng.uniform(r[0], r[1], n_tubes)
    pts = np.empty((n_tubes, n_max, 3))
    pts[:, 0] = p
    length = np.ones(n_tubes, np.int64)
    alive = np.ones(n_tubes, bool)
    for i in range(1, n_max):
        d = d + kink * rng.standard_normal((n_tubes, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p = p + step * d
        alive &= (i <= steps) & (p >= 0.0).all(axis=1) & (p <= box).all(axis=1)
        pts[:, i] = p
        length += alive
    keep = length >= 2
    pts, length, radii = pts[keep], length[keep], radii[keep]
    xyz = pts[np.arange(n_max)[None, :] < length[:, None]]
    off = np.concatenate([[0], np.cumsum(length)])
    return xyz, radii, [np.arange(off[i], off[i + 1]) for i in range(len(length))]


def spmv_bytes(n: int, nnz: int, sv: int = 8, si: int = 4) -> int:
    """Algorithmic bytes of one SSS SpMV: L(sv+si) + N(sv+si) + 2 N sv."""
    return nnz * (sv + si) + n * (sv + si) + 2 * n * sv


def pcg_iter_bytes(n: int, nnz: int, sv: int = 8, si: int = 4) -> int:
    """Algorithmic bytes of one Jacobi-PCG iteration: L(sv+si) + N(sv+si) + 13 N sv."""
    return nnz * (sv + si) + n * (sv + si) + 13 * n * sv


__all__ = ["ElectricParams", "NetworkParams", "Network", "DEASystem", "config_params",
           "generate", "largest_percolating_system", "tunnel_constants", "spmv_bytes",
           "plate_row_ranges",
           "pcg_iter_bytes", "math"]
