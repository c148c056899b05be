"""Restatement of the reference's backbone extraction -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of cicy_net_backbone.  It follows
Backbone_Network.cpp step by step, on plain Python lists, sets and dicts standing in for the
reference's HoKo->elements_cnt sets, cluster and junction vectors and structure_gnp:

  zero current                Find_zero_current                        :64-221
                              (scan: the same from voltages and node lists alone)
  same-particle junctions     Zero_current_in_same_particle_junctions_unit_resistor :225-262
  tube backbone runs          Find_backbone_cnts                       :317-409
  tube volumes, trimming      Calculate_cnt_volumes                    :412-555
                              CNT_volume_between_two_points            :557-574
  plates                      Find_backbone_gnps                       :577-661
  junctions below zero        Find_and_remove_gnp_and_mixed_junctions_below_zero_current :663-796
  dead plates                 Clear_structure_of_dead_gnps             :797-855
  tubes left without elements Remove_additional_cnts_from_cluster      :857-880
  plate volumes               Calculate_gnp_backbone_volume            :882-951
  junctions with dead ends    Remove_junctions_with_dead_particles_cnts :953-997,
                              ..._gnps :999-1048, ..._mixed :1061-1127
  plate volume inside window  Generate_Network::Approximate_gnp_volume_inside_sample
                              (Generate_Network.cpp:4750-4810)
"""
from __future__ import annotations

import math

PI = 3.1415926535897932384626433832795  # Input_Reader.h:23
ZERO = 1.0e-10                           # MathMatrix.h:18


def cnt_volume(xyz, p1, p2, r):
    length = 0.0
    for i in range(p1, p2):
        dx = xyz[i][0] - xyz[i + 1][0]
        dy = xyz[i][1] - xyz[i + 1][1]
        dz = xyz[i][2] - xyz[i + 1][2]
        length = length + math.sqrt(dx * dx + dy * dy + dz * dz)
    return length * PI * r * r


def gnp_volume(plate, box):
    c, ax, l, t = plate["centre"], plate["axes"], plate["side"], plate["thick"]
    h = 0.5 * l
    corners = []
    for s0, s1 in ((-1, -1), (1, -1), (-1, 1), (1, 1)):  # corner, right, up, up right
        corners.append([c[k] + s0 * h * ax[0][k] + s1 * h * ax[1][k] for k in range(3)])
    n = 20
    inside = 0
    for i in range(n):
        lr = i / (n - 1)
        for j in range(n):
            lu = j / (n - 1)
            ok = True
            for k in range(3):
                st = corners[0][k] + (corners[1][k] - corners[0][k]) * lr
                en = corners[2][k] + (corners[3][k] - corners[2][k]) * lr
                q = st + (en - st) * lu
                if q < 0.0 or q > box:
                    ok = False
            inside += ok
    v = l * l * t
    if inside:
        v = v * (inside / (n * n))
    return v


def scan(volt, tubes, pairs):
    """The zero current and each tube's backbone run, from voltages alone
    (Find_zero_current :64-221, Find_backbone_cnts :317-409): tubes is a list of node lists
    (one per tube, its element points in order), pairs a list of (node, node) for the other
    elements.  Returns (zero, [(idx1, idx2)] per tube)."""
    imax = ZERO
    for nodes in tubes:
        for q in range(1, len(nodes)):
            c = abs(volt[nodes[q - 1]] - volt[nodes[q]])
            if c > imax:
                imax = c
    for a, b in pairs:
        c = abs(volt[a] - volt[b])
        if c > imax:
            imax = c
    zero = imax * 1e-9
    runs = []
    for nodes in tubes:
        idx1 = idx2 = -1
        for q in range(1, len(nodes)):
            c = abs(volt[nodes[q - 1]] - volt[nodes[q]])
            if c - zero > ZERO:
                if idx1 == -1:
                    idx1 = q - 1
            elif idx1 != -1:
                idx2 = q - 1
                break
        runs.append((idx1, idx2))
    return zero, runs


def backbone(net, volt, lmm, lmm_gnp, reserved, avoid_resistance=False):
    """net: dict with the cluster's data, as read before the extraction:
      xyz [points][3], radius {tube: r}, box,
      elements {tube: sorted list of element points} (every tube that appears below),
      cluster [tubes], junctions [(p1, p2, n1, n2)] of the cluster,
      gnps [plates], structure {plate: list of points}, triangulation {plate: [(v1, v2)]},
      gjunctions [(p1, p2, g1, g2)], mixed [(p_cnt, p_gnp, cnt, gnp)], plates {g: geometry}.
    Returns (volumes dict, edited data); the lists are edited copies."""
    el = {t: list(v) for t, v in net["elements"].items()}
    cl = list(net["cluster"])
    cj = list(net["junctions"])
    cg = list(net["gnps"])
    st = {g: list(v) for g, v in net["structure"].items()}
    cgj = list(net["gjunctions"])
    cm = list(net["mixed"])

    def V(n):
        return volt[n]

    # Find_zero_current
    imax = ZERO
    currents = {}
    for t in cl:
        e = el[t]
        cur = []
        for q in range(1, len(e)):
            c = abs(V(lmm[e[q - 1]]) - V(lmm[e[q]]))
            if c > imax:
                imax = c
            cur.append(c)
        currents[t] = cur
    for p1, p2, _, _ in cj:
        c = abs(V(lmm[p1]) - V(lmm[p2]))
        if c > imax:
            imax = c
    for g in cg:
        for v1, v2 in net["triangulation"][g]:
            c = abs(V(lmm_gnp[v1]) - V(lmm_gnp[v2]))
            if c > imax:
                imax = c
    for p1, p2, _, _ in cgj:
        c = abs(V(lmm_gnp[p1]) - V(lmm_gnp[p2]))
        if c > imax:
            imax = c
    for pc, pg, _, _ in cm:
        c = abs(V(lmm[pc]) - V(lmm_gnp[pg]))
        if c > imax:
            imax = c
    zero = imax * 1e-9
    vol = dict(zero_current=zero, cnt_backbone=0.0, cnt_dead=0.0, gnp_backbone=0.0,
               gnp_dead=0.0)
    segments = {}

    # Find_backbone_cnts + Calculate_cnt_volumes, last tube first
    if cl:
        for i in range(len(cl) - 1, -1, -1):
            t = cl[i]
            cur = currents[t]
            idx1 = idx2 = -1
            for j, c in enumerate(cur):
                if c - zero > ZERO:
                    if idx1 == -1:
                        idx1 = j
                elif idx1 != -1 and idx2 == -1:
                    idx2 = j
                    break
            segments[t] = (idx1, idx2)
            e = el[t]
            r = net["radius"][t]
            ps, pe = e[0], e[-1]
            if idx1 == -1:
                vol["cnt_dead"] = vol["cnt_dead"] + cnt_volume(net["xyz"], ps, pe, r)
                if not avoid_resistance:
                    e.clear()
            else:
                p1 = e[idx1]
                vol["cnt_dead"] = vol["cnt_dead"] + cnt_volume(net["xyz"], ps, p1, r)
                if idx2 == -1:
                    vb = cnt_volume(net["xyz"], p1, pe, r)
                else:
                    p2 = e[idx2]
                    vb = cnt_volume(net["xyz"], p1, p2, r)
                    vol["cnt_dead"] = vol["cnt_dead"] + cnt_volume(net["xyz"], p2, pe, r)
                    if not avoid_resistance:
                        del e[idx2 + 1:]
                vol["cnt_backbone"] = vol["cnt_backbone"] + vb
                if not avoid_resistance and idx1 > 0:
                    del e[:idx1]
            if not avoid_resistance and not e:
                del cl[i]

    def remove_point(lst, p):
        if p in lst:
            lst.remove(p)

    if cg:
        # mixed and plate junctions below the zero current
        for i in range(len(cm) - 1, -1, -1):
            pc, pg, t, g = cm[i]
            e = el[t]
            present = pc in e
            c = abs(V(lmm[pc]) - V(lmm_gnp[pg]))
            if c - zero < ZERO or not e or not present:
                remove_point(st[g], pg)
                del cm[i]
                if present:
                    e.remove(pc)
                    if len(e) < 2:
                        e.clear()
        for i in range(len(cgj) - 1, -1, -1):
            p1, p2, g1, g2 = cgj[i]
            c = abs(V(lmm_gnp[p1]) - V(lmm_gnp[p2]))
            if c - zero < ZERO:
                remove_point(st[g1], p1)
                remove_point(st[g2], p2)
                del cgj[i]
        nb = {g: sum(1 for q in st[g] if 0 <= lmm_gnp[q] < reserved) for g in cg}

        def clear_dead():
            to_remove = 0
            for g in cg:
                nt = len(st[g])
                nj = nt - nb[g]
                if nt == 0 or (nb[g] == 0 and nj == 1) or nb[g] == nt:
                    if st[g]:
                        st[g] = []
                        to_remove += 1
                    if nj == 1:
                        to_remove += 1
                elif len(st[g]) < 2:
                    raise RuntimeError("a live plate with fewer than two points")
            return to_remove

        to_remove = clear_dead()
        while to_remove:
            for i in range(len(cgj) - 1, -1, -1):
                p1, p2, g1, g2 = cgj[i]
                if not st[g1] or not st[g2]:
                    del cgj[i]
                    if st[g1]:
                        remove_point(st[g1], p1)
                    if st[g2]:
                        remove_point(st[g2], p2)
            for i in range(len(cm) - 1, -1, -1):
                pc, pg, t, g = cm[i]
                if not el[t] or not st[g]:
                    del cm[i]
                    if st[g]:
                        remove_point(st[g], pg)
                elif pc not in el[t]:
                    del cm[i]
                    remove_point(st[g], pg)
            to_remove = clear_dead()
        for i in range(len(cl) - 1, -1, -1):
            if not el[cl[i]]:
                del cl[i]
        for g in cg:
            v = gnp_volume(net["plates"][g], net["box"])
            if st[g]:
                vol["gnp_backbone"] = vol["gnp_backbone"] + v
            else:
                vol["gnp_dead"] = vol["gnp_dead"] + v

    if not avoid_resistance:
        for i in range(len(cj) - 1, -1, -1):
            p1, p2, n1, n2 = cj[i]
            if not el[n1] or not el[n2] or p1 not in el[n1] or p2 not in el[n2]:
                del cj[i]

    return vol, dict(elements=el, cluster=cl, junctions=cj, structure=st, gjunctions=cgj,
                     mixed=cm, segments=segments)
