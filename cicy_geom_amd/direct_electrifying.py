"""Host-side mirror of the reference's Direct_Electrifying and Electrical_analysis solve
path, driving the HIP library.

Reference interface kept (rnaimehaom/CICY_geom):

=========================================  ================================================
this module                                 reference
=========================================  ================================================
DirectElectrifying.voltages / LMM_cnts      Direct_Electrifying.h:23-27 public members
DirectElectrifying.compute_voltage_field    Compute_voltage_field, Direct_Electrifying.cpp:11-113
DirectElectrifying.solve_dea_equations_cg_sss  Solve_DEA_equations_CG_SSS, :1538-1659
DirectElectrifying.update_residual_vector   Update_residual_vector, :1502-1535
DirectElectrifying.spm_v_sss                spM_V_SSS, :1678-1695
DirectElectrifying.get_voltage_vector       Get_voltage_vector, :1472-1500
calculate_parallel_resistor                 Electrical_analysis.cpp:355-473 (+ :499-605)
get_boundaries_from_direction               Electrical_analysis.cpp:477-497
vector_of_directions                        Electrical_analysis.cpp:292-346
calculate_resistances_and_resistivities     Calculate_resistances_and_resistivities, :873-932
calculate_matrix_resistance                 Calculate_matrix_resistance, :935-958
calculate_resistivity                       Calculate_resistivity, :960-987
electrical_resistance_along_each_percolated_direction   :229-289
=========================================  ================================================

Functions return 1 on success and 0 on failure and put the message in ``last_error``, as
the reference's int functions write to ``hout``.  The solve itself always runs on the GPU
(libcicy_dea.so); a missing library or device raises.
"""
from __future__ import annotations

import numpy as np

from . import dea
from .network import DEASystem, ElectricParams, Network

_CTX: dea.Context | None = None


def default_context() -> dea.Context:
    """Process-wide device context (device 0), created on first use."""
    global _CTX
    if _CTX is None:
        _CTX = dea.Context(0)
    return _CTX


def get_global_nodes(family: int) -> int:
    """Reserved (Dirichlet) node count from the family (Direct_Electrifying.cpp:115-137)."""
    if family == 6:
        return 6
    if 3 <= family <= 5:
        return 4
    if 0 <= family <= 2:
        return 2
    return -1


class DirectElectrifying:
    """One solve's object, as the reference creates one per (cluster, R_flag, family)
    (Electrical_analysis.cpp:42, :252).  `method` selects the device method; `warm_start`
    (a warmstart.WarmStartNet) replaces the ramp guess by the learned prediction, with the
    stop kept at the reference's absolute threshold tol * ||r0(ramp)||."""

    def __init__(self, ctx: dea.Context | None = None, method: int = dea.METHOD_PCG_F64,
                 warm_start=None):
        self.ctx = ctx or default_context()
        self.method = method
        self.warm_start = warm_start
        self.voltages = np.zeros(0)
        self.LMM_cnts = np.zeros(0, dtype=np.int64)   # point -> node, -1 when unmapped
        self.LMM_gnps = np.zeros(0, dtype=np.int64)
        self.points_cnt_rad: dict[int, float] = {}
        self.last_error = ""
        self.stats: dea.SolveStats | None = None
        self.system: DEASystem | None = None

    # -- Get_voltage_vector (Direct_Electrifying.cpp:1472-1500)
    def get_voltage_vector(self, reserved_nodes: int, nodes: int, r_flag: int,
                           applied_voltage: float, VEF: np.ndarray) -> int:
        if r_flag == 1:
            v = applied_voltage
        elif r_flag == 0:
            v = float(nodes)
        else:
            self.last_error = f"Error in Get_voltage_vector. The R_flag has an invalid value: {r_flag}"
            return 0
        VEF[:] = 0.0
        for i in range(1, reserved_nodes):
            VEF[i] = VEF[i - 1] + v
        return 1

    # -- spM_V_SSS on the GPU
    def spm_v_sss(self, V, rowptr, colind, diagonal, values) -> np.ndarray:
        return self.ctx.spmv(rowptr, colind, values, diagonal, V)

    # -- Update_residual_vector on the GPU: R <- R - A x0
    def update_residual_vector(self, col_ind, row_ptr, values, diagonal, V_guess, R) -> int:
        try:
            R[:] = self.ctx.update_residual(row_ptr, col_ind, values, diagonal, V_guess, R)
        except dea.DEAError as e:
            self.last_error = str(e)
            return 0
        return 1

    # -- Solve_DEA_equations_CG_SSS (the drop-in seam)
    def solve_dea_equations_cg_sss(self, nodes, reserved_nodes, tolerance, col_ind, row_ptr,
                                   values, diagonal, R, VEF, voltages_sol,
                                   max_iter: int = 0) -> int:
        """Same arguments as the reference; R and voltages_sol are updated in place
        (final recursive residual, solution) and self.voltages = [VEF | x]."""
        ok, volt, x, r, st = self.ctx.solve_dea_cg_sss(nodes, reserved_nodes, tolerance,
                                                       col_ind, row_ptr, values, diagonal, R,
                                                       VEF, voltages_sol, max_iter)
        self.stats = st
        if not ok:
            self.last_error = ("Error in Solve_DEA_equations_CG_SSS: " + self.ctx.last_error)
            return 0
        R[:] = r
        voltages_sol[:] = x
        self.voltages = volt
        return 1

    # -- Compute_voltage_field (Direct_Electrifying.cpp:11-113)
    def compute_voltage_field(self, n_cluster: int, r_flag: int, network: Network,
                              family: int, electric_param: ElectricParams | None = None,
                              tolerance: float = 1e-10, direction: int = -1) -> int:
        """Assemble the cluster's system (host C++), apply the ramp warm start when
        R_flag && R_d == 2, and solve on the GPU with this object's method."""
        if get_global_nodes(family) == -1:
            self.last_error = f"Error in Compute_voltage_field, invalid family: {family}"
            return 0
        try:
            s = network.assemble(n_cluster, r_flag, family, electric_param,
                                 direction=direction, with_xyz=self.warm_start is not None)
        except RuntimeError as e:
            self.last_error = f"Error in Compute_voltage_field: {e}"
            return 0
        self.system = s
        self.LMM_cnts = s.lmm
        self.LMM_gnps = s.lmm_gnp if s.lmm_gnp is not None else np.zeros(0, np.int64)
        R = s.r0.copy()
        x = s.x0.copy()
        if self.method == dea.METHOD_PCG_F64 and self.warm_start is None:
            return self.solve_dea_equations_cg_sss(s.nodes, s.reserved, tolerance, s.col_ind,
                                                   s.row_ptr, s.values, s.diag, R, s.vef, x)
        # the other methods on a device-resident system, same stopping rule (||r|| <=
        # tol * ||r0||, r0 = b - A x0 of the ramp guess)
        sysd = self.ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        try:
            if self.method == dea.METHOD_CONDENSED_BLOCK and s.gnp_nodes > 0:
                from .network import plate_row_ranges
                sysd.set_blocks(plate_row_ranges(network, n_cluster, s))
            sysd.set_rhs(s.r0, s.x0, s.b)
            if self.warm_start is None:
                self.stats = sysd.solve(self.method, tolerance, 0, dea.REL_R0)
            else:
                from .warmstart import Graph
                ep = electric_param or ElectricParams()
                axis = family if 0 <= family <= 2 else max(direction, 0)
                g = Graph(self.ctx, s, network.params.box, axis, ep.applied_voltage)
                try:
                    self.warm_start.warm_start(g, sysd)
                finally:
                    g.close()
                nb = float(np.linalg.norm(s.b))
                tol_b = tolerance * float(np.linalg.norm(s.r0)) / nb if nb > 0 else tolerance
                self.stats = sysd.solve(self.method, tol_b, 0, dea.REL_B)
            x, r = sysd.get()
        except dea.DEAError as e:
            self.last_error = f"Error in Compute_voltage_field: {e}"
            return 0
        finally:
            sysd.close()
        if self.stats.status not in (dea.CICY_OK, dea.CICY_MAXITER):
            self.last_error = ("Error in Compute_voltage_field: solve ended with status "
                               + self.stats.as_dict()["status_name"])
            return 0
        self.voltages = np.concatenate([s.vef, x])
        return 1


def get_boundaries_from_direction(direction: int):
    """(b1, b2) of a direction (Electrical_analysis.cpp:477-497); None if invalid."""
    return {0: (2, 4), 1: (3, 5), 2: (0, 1)}.get(direction)


def vector_of_directions(family: int, resistances=(1, 1, 1)) -> list[int] | None:
    """Directions to solve for a family (Electrical_analysis.cpp:292-346)."""
    want = {6: (0, 1, 2), 5: (1, 2), 4: (0, 2), 3: (0, 1)}
    if family in want:
        return [d for d in want[family] if resistances[d]]
    if 0 <= family <= 2:
        return [family] if resistances[family] else []
    return None


def calculate_parallel_resistor(direction: int, dea_obj: DirectElectrifying,
                                electric_param: ElectricParams | None = None,
                                ctx: dea.Context | None = None):
    """Resultant resistance of a solved cluster along `direction`
    (Electrical_analysis.cpp:355-473): I = sum over boundary elements of b1 of
    (V[n2] - V[n1]) / Re, summed in FP64 on the GPU; R = V_app / I.  Returns
    (R, I, I_check) with I_check the same sum over b2 (the reference's check)."""
    ep = electric_param or ElectricParams()
    s = dea_obj.system
    if s is None or s.direction != direction:
        raise ValueError("the solve was not assembled with boundary elements for this direction")
    ctx = ctx or dea_obj.ctx
    currents = []
    for n_bnd, n_int, re in s.boundary_elements:
        currents.append(ctx.boundary_current(n_bnd, n_int, re, dea_obj.voltages))
    i_total, i_check = currents
    return ep.applied_voltage / i_total, i_total, i_check


def window_dimensions(window) -> tuple[float, float, float]:
    """(len_x, wid_y, hei_z) of a window: a cube side (float), a 3-sequence, or an object
    with the reference's cuboid members len_x/wid_y/hei_z (Geometry_3D.h)."""
    if hasattr(window, "len_x"):
        return float(window.len_x), float(window.wid_y), float(window.hei_z)
    if np.ndim(window) == 0:
        w = float(window)
        return w, w, w
    lx, wy, hz = (float(v) for v in window)
    return lx, wy, hz


def _length_and_area(direction: int, window):
    lx, wy, hz = window_dimensions(window)
    if direction == 0:
        return lx, wy * hz
    if direction == 1:
        return wy, lx * hz
    if direction == 2:
        return hz, lx * wy
    raise ValueError(f"invalid direction {direction}: only 0, 1 or 2")


def calculate_matrix_resistance(direction: int, window, matrix_resistivity: float) -> float:
    """Resistance of the polymer alone along `direction`: rho_m * length / area
    (Calculate_matrix_resistance, Electrical_analysis.cpp:935-958)."""
    length, area = _length_and_area(direction, window)
    return matrix_resistivity * length / area


def calculate_resistivity(direction: int, window, resistance: float) -> float:
    """rho = R * area / length along `direction` (Calculate_resistivity,
    Electrical_analysis.cpp:960-987)."""
    length, area = _length_and_area(direction, window)
    return resistance * area / length


def calculate_resistance(resistors, r_matrix: float) -> float:
    """Per-direction combination (Electrical_analysis.cpp:883-917): no cluster -> the
    matrix resistance `r_matrix`, one -> that R, more -> 1/sum(1/R_k) summed in list order."""
    if len(resistors) == 0:
        return float(r_matrix)
    if len(resistors) == 1:
        return float(resistors[0])
    inv = 0.0
    for r in resistors:
        inv = inv + 1 / r
    return 1 / inv


def calculate_resistances_and_resistivities(window, electric_param: ElectricParams,
                                            parallel_resistors):
    """Resistances and resistivities along x, y, z from each direction's list of cluster
    resistances (Calculate_resistances_and_resistivities, Electrical_analysis.cpp:873-932).
    A direction without clusters takes the matrix resistance and the matrix resistivity."""
    resistors, resistivities = [0.0] * 3, [0.0] * 3
    for d in range(3):
        rk = parallel_resistors[d]
        if len(rk) == 0:
            resistivities[d] = electric_param.resistivity_matrix
            resistors[d] = calculate_matrix_resistance(d, window,
                                                       electric_param.resistivity_matrix)
        else:
            resistors[d] = calculate_resistance(rk, 0.0)
            resistivities[d] = calculate_resistivity(d, window, resistors[d])
    return resistors, resistivities


def electrical_resistance_along_each_percolated_direction(
        network: Network, n_cluster: int, family: int,
        electric_param: ElectricParams | None = None, tolerance: float = 1e-10,
        resistances=(1, 1, 1), method: int = dea.METHOD_PCG_F64,
        ctx: dea.Context | None = None, warm_start=None):
    """One R_flag=1 solve per percolated direction with family := direction
    (Electrical_analysis.cpp:229-289).  Returns {direction: (R, I, I_check, stats)}."""
    ep = electric_param or ElectricParams()
    dirs = vector_of_directions(family, resistances)
    if dirs is None:
        raise ValueError(f"Invalid family: {family}")
    out = {}
    for d in dirs:
        de = DirectElectrifying(ctx, method, warm_start)
        if not de.compute_voltage_field(n_cluster, 1, network, d, ep, tolerance, direction=d):
            raise RuntimeError(de.last_error)
        R, I, Ic = calculate_parallel_resistor(d, de, ep)
        out[d] = (R, I, Ic, de.stats)
    return out


def network_resistances(network: Network, electric_param: ElectricParams | None = None,
                        tolerance: float = 1e-10, method: int = dea.METHOD_PCG_F64,
                        ctx: dea.Context | None = None, with_details: bool = False,
                        warm_start=None):
    """Resistance of the network along x, y, z: every percolating cluster solved per
    direction and combined in parallel, a direction without one taking the matrix
    resistance of the window (Perform_analysis_on_clusters, Electrical_analysis.cpp:10-145,
    without the backbone step; then Calculate_resistances_and_resistivities, :873-932).

    Returns [R_x, R_y, R_z]; with `with_details`, (resistances, resistivities,
    {direction: [(cluster, R, I, I_check, stats), ...]})."""
    ep = electric_param or ElectricParams()
    per_dir = {0: [], 1: [], 2: []}
    details = {0: [], 1: [], 2: []}
    for k in range(network.count("clusters")):
        _, fam = network.cluster(k)
        if fam < 0:
            continue
        res = electrical_resistance_along_each_percolated_direction(
            network, k, fam, ep, tolerance, method=method, ctx=ctx, warm_start=warm_start)
        for d, (R, I, Ic, st) in res.items():
            per_dir[d].append(R)
            details[d].append((k, R, I, Ic, st))
    resistors, resistivities = calculate_resistances_and_resistivities(
        network.params.box, ep, [per_dir[d] for d in range(3)])
    if with_details:
        return resistors, resistivities, details
    return resistors


def network_resistances_device(network: Network, electric_param: ElectricParams | None = None,
                               tolerance: float = 1e-10,
                               method: int = dea.METHOD_CONDENSED_BLOCK,
                               ctx: dea.Context | None = None, with_details: bool = False):
    """network_resistances with the network copied to the device once and every (cluster,
    direction) system assembled there (cicy_dnet_assemble: the assembly half of
    Compute_voltage_field, Direct_Electrifying.cpp:11-113) and handed to the solver without
    a copy through the host; the plate groups of the block method come with the assembly.
    The stopping rule (||r|| <= tol * ||r0|| from the ramp guess), the currents and the
    combination are network_resistances'.  Returns what network_resistances returns."""
    ep = electric_param or ElectricParams()
    ctx = ctx or default_context()
    per_dir = {0: [], 1: [], 2: []}
    details = {0: [], 1: [], 2: []}
    dn = network.device(ctx)
    try:
        for k in range(network.count("clusters")):
            _, fam = network.cluster(k)
            if fam < 0:
                continue
            dirs = vector_of_directions(fam)
            if dirs is None:
                raise ValueError(f"Invalid family: {fam}")
            for d in dirs:
                a = dn.assemble(k, 1, d, ep, d)
                try:
                    sysd = a.system()
                    try:
                        if method == dea.METHOD_CONDENSED_BLOCK and a.gnp_nodes > 0:
                            sysd.set_blocks(a.blocks)
                        st = sysd.solve(method, tolerance, 0, dea.REL_R0)
                        if st.status not in (dea.CICY_OK, dea.CICY_MAXITER):
                            raise RuntimeError(f"solve of cluster {k}, direction {d}: "
                                               + st.as_dict()["status_name"])
                        x, _ = sysd.get()
                    finally:
                        sysd.close()
                    volt = np.concatenate([a.vef, x])
                    i1, i2 = (ctx.boundary_current(*e, volt) for e in a.boundary_elements())
                finally:
                    a.close()
                R = ep.applied_voltage / i1
                per_dir[d].append(R)
                details[d].append((k, R, i1, i2, st))
    finally:
        dn.close()
    resistors, resistivities = calculate_resistances_and_resistivities(
        network.params.box, ep, [per_dir[d] for d in range(3)])
    if with_details:
        return resistors, resistivities, details
    return resistors


def network_resistances_batched(network: Network,
                                electric_param: ElectricParams | None = None,
                                tolerance: float = 1e-10,
                                method: int = dea.METHOD_PCG_F64,
                                ctx: dea.Context | None = None, with_details: bool = False):
    """network_resistances with every (cluster, direction) system solved together in one
    batch (cicy_batch: concatenated systems, per-system scalars and stopping tests in shared
    launches) instead of one after another as the reference's cluster and direction loops do
    (Electrical_analysis.cpp:34-105, :249-289).  Each system starts from its ramp guess and
    stops by the reference's rule ||r|| <= tol * ||r0||; currents and the combination are as
    in network_resistances.  `method` is METHOD_PCG_F64 or METHOD_MIXED."""
    ep = electric_param or ElectricParams()
    ctx = ctx or default_context()
    jobs = []
    for k in range(network.count("clusters")):
        _, fam = network.cluster(k)
        if fam < 0:
            continue
        dirs = vector_of_directions(fam)
        if dirs is None:
            raise ValueError(f"Invalid family: {fam}")
        for d in dirs:
            jobs.append((k, d, network.assemble(k, 1, d, ep, direction=d)))
    per_dir = {0: [], 1: [], 2: []}
    details = {0: [], 1: [], 2: []}
    stats = []
    if jobs:
        B = ctx.batch([(s.row_ptr, s.col_ind, s.values, s.diag) for _, _, s in jobs])
        try:
            for i, (_, _, s) in enumerate(jobs):
                B.set_rhs(i, s.r0, s.x0, s.b)
            stats, _ = B.solve(method, tolerance, 0, dea.REL_R0)
            for i, (k, d, s) in enumerate(jobs):
                st = stats[i]
                if st.status not in (dea.CICY_OK, dea.CICY_MAXITER):
                    raise RuntimeError(f"batched solve of cluster {k}, direction {d}: "
                                       + st.as_dict()["status_name"])
                x, _ = B.get(i)
                volt = np.concatenate([s.vef, x])
                i1 = ctx.boundary_current(*s.boundary_elements[0], volt)
                i2 = ctx.boundary_current(*s.boundary_elements[1], volt)
                R = ep.applied_voltage / i1
                per_dir[d].append(R)
                details[d].append((k, R, i1, i2, st))
        finally:
            B.close()
    resistors, resistivities = calculate_resistances_and_resistivities(
        network.params.box, ep, [per_dir[d] for d in range(3)])
    if with_details:
        return resistors, resistivities, details
    return resistors
