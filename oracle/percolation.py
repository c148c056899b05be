"""Restatement of the reference's tube clusters and percolation test -- TEST INFRASTRUCTURE
ONLY.  Only tests/ may import this module, as the checker of cicy_cnt_percolation.  It follows
Hoshen_Kopelman.cpp, on Python lists, sets and dicts:
  Make_particle_clusters :2331-2364, Find_percolated_clusters :2366-2396,
  Find_clusters_connected_to_boundaries :2398-2468, Add_clusters_in_boundary :2470-2495,
  Add_percolated_direction :2497-2522, Determine_family_of_percolated_clusters :2524-2589,
  Family_map :2591-2622, Group_junctions_same_particle :2672-2705."""

FAMILY_MAP = {1: 0, 11: 1, 21: 2, 12: 3, 22: 4, 32: 5, 33: 6}


def cnt_percolation(labels, n_clusters, boundary_cnt, junctions, cnts_inside):
    """Returns (clusters_cnt_tmp, isolated_cnt, percolated_labels {label: index}, family,
    clusters_cnt, cluster_cnt_junctions)."""
    clusters_tmp = [[] for _ in range(n_clusters)]
    isolated = [[]]
    for t in cnts_inside:
        L = labels[t]
        if L != -1:
            clusters_tmp[L].append(t)
        else:
            isolated[-1].append(t)
    dirs = [set() for _ in range(n_clusters)]
    for d, (b1, b2) in enumerate(((2, 4), (3, 5), (0, 1))):
        boundary1 = {labels[t] for t in boundary_cnt[b1] if labels[t] != -1}
        if boundary1:
            for t in boundary_cnt[b2]:
                L = labels[t]
                if L in boundary1:
                    dirs[L].add(10 * d + 1)
    percolated, family, clusters = {}, [], []
    for i in range(n_clusters):
        if not dirs[i]:
            isolated.append(clusters_tmp[i])
        else:
            percolated[i] = len(clusters)
            family.append(FAMILY_MAP[sum(dirs[i])])
            clusters.append(clusters_tmp[i])
    groups = [[] for _ in clusters]
    for q, (_, a, _, b) in enumerate(junctions):
        L1, L2 = labels[a], labels[b]
        if L1 != L2:
            raise RuntimeError("a junction between tubes of different clusters")
        if L1 in percolated:
            groups[percolated[L1]].append(q)
    return clusters_tmp, isolated, percolated, family, clusters, groups
