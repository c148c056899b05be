"""Restatement of the reference's tube contact search and Hoshen-Kopelman labelling --
TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of cicy_cnt_contacts.  It follows, on
Python lists, dicts and sets standing in for the reference's vectors, maps and sets:

  subregions          Contact_grid::Generate_contact_grid (tube part)  Contact_grid.cpp:18-80
                      Adjust_regions_if_needed                           :82-102
                      Fill_sectioned_domain_cnts                         :104-161
                      Calculate_region_coordinates                       :163-184
                      Calculate_overlapping_flags                        :186-212
                      Assign_point_to_regions_cnts, Calculate_t          :214-295
  contacts, labels    Hoshen_Kopelman::Label_cnts_in_window   Hoshen_Kopelman.cpp:249-371
                      HK76 :2908-2953, Find_root :2955-2967, Merge_labels :2969-2989
                      Cleanup_labels                          :374-425
  junctions           Compress_cnt_cnt_contact_segments       :443-555
  elements            Complete_cnt_elements                   :557-573
                      (Make_cnt_clusters, :203-247, runs the four in this order)

The distances of one subregion are computed with numpy for speed: the same IEEE operations
in the same order (dx*dx + dy*dy + dz*dz, then sqrt, then - r1 - r2), no fused multiply-add.
The contacts are then visited one by one in the reference's (subregion, j, k) order."""
from __future__ import annotations

import numpy as np

ZERO = 1.0e-10  # MathMatrix.h:18


def subregions(window_min, window_len, gs_min, overlap_cnt, overlap_gnp, cnts_inside, structure,
               xyz):
    """sectioned_domain_cnt: list of point-id lists, one per subregion, and (n, l)."""
    n = [max(1, int(window_len[k] / gs_min[k])) for k in range(3)]
    l = [float(g) for g in gs_min]
    ov = max(overlap_cnt, overlap_gnp)
    for k in range(3):
        if gs_min[k] < 2 * ov:
            l[k] = 2.0 * (ov + ZERO)
            n[k] = 1 + int(window_len[k] / l[k])
    regions = [[] for _ in range(n[0] * n[1] * n[2])]

    def t_of(a, b, c):
        return a + b * n[0] + c * n[0] * n[1]

    for cnt in cnts_inside:
        for P in structure[cnt]:
            p = xyz[P]
            co = []
            for k in range(3):
                a = int((p[k] - window_min[k]) / l[k])
                if a < 0:
                    a = 0
                elif a >= n[k]:
                    a = n[k] - 1
                co.append(a)
            f = [0, 0, 0]
            for k in range(3):
                x1 = float(co[k]) * l[k] + window_min[k]
                x2 = x1 + l[k]
                if co[k] > 0 and p[k] >= x1 and p[k] <= x1 + overlap_cnt:
                    f[k] = -1
                elif co[k] < n[k] - 1 and p[k] >= x2 - overlap_cnt and p[k] <= x2:
                    f[k] = 1
            regions[t_of(*co)].append(P)
            for k in range(3):
                if f[k]:
                    c2 = list(co)
                    c2[k] += f[k]
                    regions[t_of(*c2)].append(P)
    return regions, n, l


def find_root(L, ll):
    r = L
    while ll[r] <= 0:
        r = -ll[r]
        if ll[r] == 0:
            return 0
    return r


def hk76(c1, c2, state, labels, ll):
    if labels[c1] == -1 and labels[c2] == -1:
        labels[c1] = labels[c2] = state[0]
        state[0] += 1
        ll.append(2)
    elif labels[c1] == -1:
        L = labels[c2]
        if ll[L] <= 0:
            L = find_root(L, ll)
        labels[c1] = L
        ll[L] += 1
    elif labels[c2] == -1:
        L = labels[c1]
        if ll[L] <= 0:
            L = find_root(L, ll)
        labels[c2] = L
        ll[L] += 1
    elif labels[c1] != labels[c2]:
        r1, r2 = find_root(labels[c1], ll), find_root(labels[c2], ll)
        if ll[r1] <= 0 or ll[r2] <= 0:
            raise RuntimeError("merging a label that is not a root")
        if r1 < r2:
            ll[r1] += ll[r2]
            ll[r2] = -r1
        elif r2 < r1:
            ll[r2] += ll[r1]
            ll[r1] = -r2


def contact_pairs(regions, xyz, flag, radii, vdw, tunnel):
    """Every tunnelling pair in the reference's visiting order:
    [(region, j, k, P1, P2, dist_junc)]."""
    out = []
    X = np.asarray(xyz, np.float64)
    for i, reg in enumerate(regions):
        m = len(reg)
        if m < 2:
            continue
        idx = np.asarray(reg, np.int64)
        p = X[idx]
        c = np.asarray([flag[P] for P in reg], np.int64)
        r = np.asarray(radii, np.float64)[c]
        dx = p[:, None, 0] - p[None, :, 0]
        dy = p[:, None, 1] - p[None, :, 1]
        dz = p[:, None, 2] - p[None, :, 2]
        dist = np.sqrt(dx * dx + dy * dy + dz * dz)
        dj = dist - r[:, None] - r[None, :]
        dj = np.where(dj < vdw, vdw, dj)
        ok = (dj <= tunnel) & (c[:, None] != c[None, :]) & np.triu(np.ones((m, m), bool), 1)
        for j, k in zip(*np.nonzero(ok)):  # row-major: j, then k
            out.append((i, int(j), int(k), reg[j], reg[k], float(dj[j, k])))
    return out


def cnt_clusters(window_min, window_len, gs_min, overlap_cnt, overlap_gnp, cutoffs, radii,
                 structure, xyz, cnts_inside=None):
    """cutoffs = (van_der_Waals_dist, tunneling_dist, min_points).  Returns dict(labels,
    n_labels, junctions [(Pi, CNTi, Pj, CNTj, d)], elements [sorted point lists per tube],
    contacts (the visiting-order list), regions)."""
    vdw, tunnel, min_points = cutoffs
    n_cnts = len(structure)
    if cnts_inside is None:
        cnts_inside = list(range(n_cnts))
    flag = {}
    for i, s in enumerate(structure):
        for P in s:
            flag[int(P)] = i
    regions, _, _ = subregions(window_min, window_len, gs_min, overlap_cnt, overlap_gnp,
                               cnts_inside, structure, xyz)
    contacts = contact_pairs(regions, xyz, flag, radii, vdw, tunnel)
    labels = [-1] * n_cnts
    ll = []
    state = [0]
    contact_elements = [dict() for _ in range(n_cnts)]
    pc, pcd = {}, {}
    for _, _, _, P1, P2, dj in contacts:
        c1, c2 = flag[P1], flag[P2]
        hk76(c1, c2, state, labels, ll)
        a, b = min(c1, c2), max(c1, c2)
        Pa = P1 if c1 < c2 else P2
        Pb = P1 if c2 < c1 else P2
        contact_elements[a].setdefault(b, set()).add(Pa)
        key = (Pa, b)
        if key not in pc or pcd[key] > dj:
            pcd[key] = dj
            pc[key] = Pb
    # Cleanup_labels
    label_map = [-1] * len(ll)
    proper = 0
    for i in range(len(ll)):
        if ll[i] > 0:
            label_map[i] = proper
            proper += 1
    for i in range(n_cnts):
        if labels[i] != -1:
            labels[i] = label_map[find_root(labels[i], ll)]
    # Compress_cnt_cnt_contact_segments
    elements = [set() for _ in range(n_cnts)]
    junctions = []
    for i in range(n_cnts):
        for cj in sorted(contact_elements[i]):
            pts = sorted(contact_elements[i][cj])
            Pi0 = pts[0]
            Pj0 = pc[(Pi0, cj)]
            dmin = pcd[(Pi0, cj)]
            Pi_j, Pj_j = Pi0, Pj0
            for Pi1 in pts[1:]:
                Pj1 = pc[(Pi1, cj)]
                d = pcd[(Pi1, cj)]
                if Pi1 - Pi0 >= min_points or abs(Pj1 - Pj0) >= min_points:
                    junctions.append((Pi_j, i, Pj_j, cj, dmin))
                    elements[i].add(Pi_j)
                    elements[cj].add(Pj_j)
                    dmin, Pi_j, Pj_j = d, Pi1, Pj1
                elif d < dmin:
                    dmin, Pi_j, Pj_j = d, Pi1, Pj1
                Pi0, Pj0 = Pi1, Pj1
            junctions.append((Pi_j, i, Pj_j, cj, dmin))
            elements[i].add(Pi_j)
            elements[cj].add(Pj_j)
    # Complete_cnt_elements
    for i in range(n_cnts):
        if elements[i]:
            elements[i].add(int(structure[i][0]))
            elements[i].add(int(structure[i][-1]))
    return dict(labels=labels, n_labels=proper, junctions=junctions,
                elements=[sorted(e) for e in elements], contacts=contacts, regions=regions)
