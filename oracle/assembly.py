"""Dictionary-based restatement of the reference's DEA system assembly.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Pure-Python loops, for small
networks.  Each step follows rnaimehaom/CICY_geom's Direct_Electrifying.cpp with a
Python dict standing in for each row's std::map<long,double>:

  reserved_nodes           Get_global_nodes              :115-137
  boundaries_of_family     Get_vector_of_boundaries      :238-305
  LMM                      Map_points_at_boundaries :217-236, LM_matrix_for_cnts :139-215
  CNT elements             Fill_2d_matrices_cnts :489-553, Three_points_in_element :555-608
  add_new / add_existing   Add_new_elements_to_2d_sparse_matrix :648-668,
                           Add_to_existing_elements_in_2d_sparse_matrix :670-694
  junctions                Fill_2d_matrices_cnt_junctions :738-852
  plate nodes              LM_matrix_for_gnps :307-349
  tube-plate junctions     Fill_2d_matrices_mixed_junctions :903-1019
  plate triangulation      Fill_2d_matrices_gnp :1021-1196, Calculate_resistance_gnp :1198-1232
  plate-plate junctions    Fill_2d_matrices_gnp_junctions :1234-1327
  SSS + RHS                From_2d_to_1d_vectors :1329-1394, Get_voltage_vector :1472-1500
  ramp guess               Initial_guess_for_CG / Initial_guess_for_cnt_nodes :1396-1470
  boundary elements        Electrical_analysis.cpp:499-605 (tubes), :607-675 (plates)

The plate triangulation itself is an input here (it is pinned separately against
scipy's Delaunay triangulation in tests/test_gnp.py).
"""
from __future__ import annotations

import math

import numpy as np

PI = 3.1415926535897932384626433832795
ZERO = 1.0e-10


def reserved_nodes(family):
    if family == 6:
        return 6
    if 3 <= family <= 5:
        return 4
    if 0 <= family <= 2:
        return 2
    return -1


def boundaries_of_family(family):
    return {0: [2, 4], 1: [3, 5], 2: [0, 1], 3: [2, 4, 3, 5], 4: [2, 4, 0, 1],
            5: [3, 5, 0, 1], 6: [2, 4, 3, 5, 0, 1]}[family]


def resistance_cnt(xyz, p1, p2, radius, rho):
    length = 0.0
    for i in range(p1, p2):
        dx = xyz[i][0] - xyz[i + 1][0]
        dy = xyz[i][1] - xyz[i + 1][1]
        dz = xyz[i][2] - xyz[i + 1][2]
        length = length + math.sqrt(dx * dx + dy * dy + dz * dz)
    return rho * length / (PI * radius * radius)


def junction_resistance(dist, r1, r2, elec, cnt1=True, cnt2=True):
    """Calculate_junction_resistance (Direct_Electrifying.cpp:854-901): a tube's radius is
    doubled to its diameter, a plate's thickness is used as is."""
    if not elec["junction_exponential"]:
        return elec["junction_resistance"]
    jl1 = r1 + r1 if cnt1 else r1
    jl2 = r2 + r2 if cnt2 else r2
    a = jl1 * jl1 if (jl1 - jl2 < ZERO) else jl2 * jl2
    return elec["c1"] * dist * math.exp(elec["c2"] * dist) / a


def resistance_gnp(p1, p2, rad1, rad2, elec):
    """Calculate_resistance_gnp (Direct_Electrifying.cpp:1198-1232): the resistivity
    tensor diag(rho_s, rho_s, rho_t) is taken in the global axes."""
    dx, dy, dz = p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]
    length = math.sqrt(dx * dx + dy * dy + dz * dz)
    if length < ZERO:
        raise ValueError("triangulation edge of zero length")
    ux, uy, uz = (p2[0] - p1[0]) / length, (p2[1] - p1[1]) / length, (p2[2] - p1[2]) / length
    rx, ry, rz = ux * elec["rho_gnp_surf"], uy * elec["rho_gnp_surf"], uz * elec["rho_gnp_t"]
    rho = rx * ux + ry * uy + rz * uz
    return rho * length / (PI * rad1 * rad2)


def assemble(family, r_flag, boundary_pts, cluster_tubes, elements, xyz, radii, junctions,
             cluster_junctions, elec, box, gnp=None):
    """Returns a dict with the SSS system, b, VEF, LMM (dict), ramp guess and r0.

    `gnp` (optional) describes the cluster's plates: boundary_pts[6] (plate points at each
    face), cluster_gnps, structure {g: points}, xyz (plate points), thick {g: t},
    triangulation {g: [(v1, v2), ...]}, mixed (all tube-plate junctions as (p_cnt, p_gnp,
    cnt, gnp, dist)), cluster_mixed, gjunc (all plate-plate junctions as (p1, p2, g1, g2,
    dist)), cluster_gjunc.  Plate nodes come after the tube nodes."""
    rd = reserved_nodes(family)
    lmm = {}
    for n, b in enumerate(boundaries_of_family(family)):
        for p in boundary_pts[b]:
            lmm[int(p)] = n
    gn = rd
    for t in cluster_tubes:
        e = [int(p) for p in elements[t]]
        if len(e) <= 1:
            raise ValueError("tube with not enough points to form an element")
        if e[0] not in lmm:
            lmm[e[0]] = gn
            gn += 1
        for p in e[1:-1]:
            lmm[p] = gn
            gn += 1
        if e[-1] not in lmm:
            lmm[e[-1]] = gn
            gn += 1
    lmm_g = {}
    if gnp is not None:
        for n, b in enumerate(boundaries_of_family(family)):
            for p in gnp["boundary_pts"][b]:
                lmm_g[int(p)] = n
        for g in gnp["cluster_gnps"]:
            for p in gnp["structure"][g]:
                if int(p) not in lmm_g:
                    lmm_g[int(p)] = gn
                    gn += 1
    nodes = gn
    col_values = [dict() for _ in range(nodes)]
    diagonal = [0.0] * nodes

    def add_new(n1, n2, g):
        diagonal[n1] += g
        diagonal[n2] += g
        if n1 > n2:
            col_values[n1][n2] = -g
        else:
            col_values[n2][n1] = -g

    def add_existing(n1, n2, g):
        diagonal[n1] += g
        diagonal[n2] += g
        if n1 > n2:
            col_values[n1][n2] = col_values[n1].get(n2, 0.0) - g
        else:
            col_values[n2][n1] = col_values[n2].get(n1, 0.0) - g

    for t in cluster_tubes:
        e = [int(p) for p in elements[t]]
        if len(e) == 3 and lmm[e[0]] == lmm[e[-1]]:
            g1 = g2 = 1.0
            if r_flag == 1:
                g1 = 1 / resistance_cnt(xyz, e[0], e[1], radii[t], elec["rho"])
                g2 = 1 / resistance_cnt(xyz, e[1], e[2], radii[t], elec["rho"])
            n1, n2 = lmm[e[0]], lmm[e[1]]
            diagonal[n1] += g1 + g2
            diagonal[n2] += g1 + g2
            col_values[n2][n1] = -(g1 + g2)
            continue
        p1 = e[0]
        for p2 in e[1:]:
            re = resistance_cnt(xyz, p1, p2, radii[t], elec["rho"]) if r_flag == 1 else 1.0
            g = 1 / re if r_flag == 1 else re
            add_new(lmm[p1], lmm[p2], g)
            p1 = p2
    for j in cluster_junctions:
        P1, P2, N1, N2, dist = junctions[j]
        n1 = lmm.get(int(P1), -1)
        n2 = lmm.get(int(P2), -1)
        if n1 == -1 or n2 == -1:
            continue
        g = 1.0
        if r_flag == 1:
            g = 1 / junction_resistance(dist, radii[N1], radii[N2], elec)
        if n1 >= rd and n2 >= rd:
            add_new(n1, n2, g)
        elif rd > 2:
            add_existing(n1, n2, g)
    points_cnt_rad = {}
    if gnp is not None:
        gx = gnp["xyz"]
        for j in gnp["cluster_mixed"]:
            p_cnt, p_gnp, cnt, gp, dist = gnp["mixed"][j]
            n1 = lmm.get(int(p_cnt), -1)
            n2 = lmm_g.get(int(p_gnp), -1)
            if n1 == -1 or n2 == -1:
                continue
            g = 1.0
            if r_flag == 1:
                g = 1 / junction_resistance(dist, radii[cnt], gnp["thick"][gp], elec,
                                            cnt1=True, cnt2=False)
            if n1 >= rd and n2 >= rd:
                add_new(n1, n2, g)
            elif rd > 2:
                add_existing(n1, n2, g)
            points_cnt_rad[int(p_gnp)] = radii[cnt]
        for gp in gnp["cluster_gnps"]:
            for v1, v2 in reversed(gnp["triangulation"][gp]):
                node1, node2 = lmm_g[int(v1)], lmm_g[int(v2)]
                g = 1.0
                if r_flag == 1:
                    rad1 = points_cnt_rad.get(int(v1), gnp["thick"][gp] / 2)
                    rad2 = points_cnt_rad.get(int(v2), gnp["thick"][gp] / 2)
                    g = 1 / resistance_gnp(gx[int(v1)], gx[int(v2)], rad1, rad2, elec)
                add_new(node1, node2, g)
        for j in gnp["cluster_gjunc"]:
            p1, p2, g1, g2, dist = gnp["gjunc"][j]
            n1 = lmm_g.get(int(p1), -1)
            n2 = lmm_g.get(int(p2), -1)
            if n1 == -1 or n2 == -1:
                continue
            g = 1.0
            if r_flag == 1:
                g = 1 / junction_resistance(dist, gnp["thick"][g1], gnp["thick"][g2], elec,
                                            cnt1=False, cnt2=False)
            add_new(n1, n2, g)
    vmag = elec["applied_voltage"] if r_flag == 1 else float(nodes)
    vef = [0.0] * rd
    for i in range(1, rd):
        vef[i] = vef[i - 1] + vmag
    n = nodes - rd
    R = [0.0] * n
    row_ptr = [0]
    col_ind, values = [], []
    for i in range(rd, nodes):
        for col in sorted(col_values[i]):
            kij = col_values[i][col]
            if col >= rd:
                values.append(kij)
                col_ind.append(col - rd)
            else:
                R[i - rd] = R[i - rd] - (kij * vef[col])
        row_ptr.append(len(values))
    diag = diagonal[rd:]
    if any(abs(d) < ZERO for d in diag):
        raise ValueError("zero element on the diagonal")
    x0 = ramp_guess(n, rd, r_flag, family, cluster_tubes, elements, lmm, xyz,
                    elec["applied_voltage"], box)
    return dict(nodes=nodes, reserved=rd, row_ptr=np.array(row_ptr, np.int64),
                col_ind=np.array(col_ind, np.int64), values=np.array(values),
                diag=np.array(diag), b=np.array(R), vef=np.array(vef), lmm=lmm, x0=x0,
                lmm_gnp=lmm_g, points_cnt_rad=points_cnt_rad)


def ramp_guess(n, rd, r_flag, family, cluster_tubes, elements, lmm, xyz, applied_voltage,
               box):
    """Initial guess of the free nodes (Direct_Electrifying.cpp:80-95 with
    Initial_guess_for_CG / Select_coordinate / Initial_guess_for_cnt_nodes, :1396-1470).

    Only when R_flag == 1 and there are two reserved nodes: for every element point of every
    cluster tube whose node is not reserved (node >= 2), x0[node - rd] = V_app / L_window *
    (max_coord - coord) along the family's axis (0 x, 1 y, 2 z).  The window is [0, box]^3,
    so L_window = max_coord = box.  Plate nodes keep 0."""
    x0 = np.zeros(n)
    if r_flag == 1 and rd == 2:
        vpc = applied_voltage / box
        ax = family
        for t in cluster_tubes:
            for p in elements[t]:
                nd = lmm[int(p)]
                if nd >= 2:
                    x0[nd - rd] = vpc * (box - xyz[int(p)][ax])
    return x0


def boundary_elements(direction, boundary_tubes, elements, lmm, xyz, radii, rho, gnp=None,
                      lmm_gnp=None, points_cnt_rad=None, elec=None):
    """(n_bnd, n_int, re) for b1 and b2 of `direction` (Electrical_analysis.cpp:499-605),
    followed by the plate triangulation edges at the boundary (:607-675) when `gnp` is
    given (boundary_gnps[6] plus the keys of assemble's `gnp`)."""
    b1, b2 = {0: (2, 4), 1: (3, 5), 2: (0, 1)}[direction]
    out = []
    for node, b in enumerate((b1, b2)):
        nb, ni, rr = [], [], []
        for t in boundary_tubes[b]:
            e = [int(p) for p in elements[t]]
            if len(e) < 2:
                continue
            for p1, p2 in ((e[0], e[1]), (e[-1], e[-2])):
                if p1 not in lmm:
                    continue
                n1 = lmm[p1]
                if n1 > 1:
                    continue
                if p2 not in lmm:
                    continue
                n2 = lmm[p2]
                lo, hi = (p2, p1) if p1 > p2 else (p1, p2)
                nb.append(n1)
                ni.append(n2)
                rr.append(resistance_cnt(xyz, lo, hi, radii[t], rho))
        if gnp is not None:
            in_cluster = set(int(g) for g in gnp["cluster_gnps"])
            for gp in gnp["boundary_gnps"][b]:
                if int(gp) not in in_cluster:
                    continue
                for v1, v2 in gnp["triangulation"][int(gp)]:
                    na, nb_ = lmm_gnp[int(v1)], lmm_gnp[int(v2)]
                    if na > 1 and nb_ > 1:
                        continue
                    if na == nb_:
                        raise ValueError("both nodes at the same boundary")
                    n1, n2 = (na, nb_) if na == node else (nb_, na)
                    t2 = gnp["thick"][int(gp)] / 2
                    rad1 = points_cnt_rad.get(int(v1), t2)
                    rad2 = points_cnt_rad.get(int(v2), t2)
                    nb.append(n1)
                    ni.append(n2)
                    rr.append(resistance_gnp(gnp["xyz"][int(v1)], gnp["xyz"][int(v2)], rad1,
                                             rad2, elec))
        out.append((np.array(nb, np.int64), np.array(ni, np.int64), np.array(rr)))
    return out


def build_from_edges(n_nodes, reserved, edges, vef):
    """SSS system of a hand-built resistor network: nodes 0..reserved-1 are Dirichlet
    nodes with voltages vef; edges (i, j, g) are added with Add_new semantics (the
    diagonal accumulates, the off-diagonal is overwritten)."""
    col_values = [dict() for _ in range(n_nodes)]
    diagonal = [0.0] * n_nodes
    for i, j, g in edges:
        diagonal[i] += g
        diagonal[j] += g
        if i > j:
            col_values[i][j] = -g
        else:
            col_values[j][i] = -g
    n = n_nodes - reserved
    R = [0.0] * n
    row_ptr = [0]
    col_ind, values = [], []
    for i in range(reserved, n_nodes):
        for col in sorted(col_values[i]):
            kij = col_values[i][col]
            if col >= reserved:
                values.append(kij)
                col_ind.append(col - reserved)
            else:
                R[i - reserved] = R[i - reserved] - kij * vef[col]
        row_ptr.append(len(values))
    return dict(nodes=n_nodes, reserved=reserved, row_ptr=np.array(row_ptr, np.int64),
                col_ind=np.array(col_ind, np.int64), values=np.array(values),
                diag=np.array(diagonal[reserved:]), b=np.array(R), vef=np.array(vef, float))
