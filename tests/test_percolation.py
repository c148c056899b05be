"""Tube clusters and percolation (cicy_cnt_percolation) against oracle/percolation.py, and the
restatement pinned independently: clusters are the connected components of the junction graph
(scipy), a cluster percolates along an axis exactly when it has tubes on both of that axis'
faces (found from the tube end points), and families follow Family_map's table
(Hoshen_Kopelman.cpp:2591-2622: x 0, y 1, z 2, xy 3, xz 4, yz 5, xyz 6).  No GPU: labels and
junctions come from oracle.hk."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from cicy_geom_amd import clusters, network
from oracle import hk
from oracle import percolation as op

import test_hk_oracle as th

FAMILY = {frozenset("x"): 0, frozenset("y"): 1, frozenset("z"): 2, frozenset("xy"): 3,
          frozenset("xz"): 4, frozenset("yz"): 5, frozenset("xyz"): 6}
# boundary numbers of the reference's window cut: x_min 4, x_max 2, y_min 5, y_max 3,
# z_min 1, z_max 0
FACES = {4: (0, 0), 2: (0, 1), 5: (1, 0), 3: (1, 1), 1: (2, 0), 0: (2, 1)}


def boundary_lists(xyz, structure, box, eps):
    out = [[] for _ in range(6)]
    for t, s in enumerate(structure):
        for P in (s[0], s[-1]):
            for b, (ax, hi) in FACES.items():
                c = xyz[P][ax]
                if (hi and c >= box - eps) or (not hi and c <= eps):
                    if t not in out[b]:
                        out[b].append(t)
    return out


@pytest.fixture(scope="module", params=[(21, 700, 0.5), (22, 500, 0.4)])
def case(request):
    seed, n, box = request.param
    xyz, radii, structure = network.wavy_tubes(seed, n, box, n_min=30, n_max=90)
    res = th.run(xyz, radii, structure, box)
    return xyz, structure, box, res, boundary_lists(xyz, structure, box, 0.011)


def test_matches_the_restatement(case):
    xyz, structure, box, res, bl = case
    lab, nl, junc = res["labels"], res["n_labels"], [j[:4] for j in res["junctions"]]
    got = clusters.cnt_percolation(lab, nl, bl, np.array(junc, np.int64).reshape(-1, 4))
    tmp, iso, perc, fam, cl, groups = op.cnt_percolation(lab, nl, bl, junc,
                                                         list(range(len(structure))))
    assert [c.tolist() for c in got["clusters"]] == tmp
    assert got["isolated"].tolist() == iso[0]
    assert got["n_perc"] == len(perc) > 0
    assert {L: int(got["perc_index"][L]) for L in range(nl) if got["perc_index"][L] >= 0} == perc
    assert [int(got["family"][L]) for L in sorted(perc)] == fam
    want_jc = [-1] * len(junc)
    for k, g in enumerate(groups):
        for q in g:
            want_jc[q] = k
    assert got["junction_cluster"].tolist() == want_jc


def test_restatement_against_components_and_faces(case):
    xyz, structure, box, res, bl = case
    lab, nl = res["labels"], res["n_labels"]
    junc = [j[:4] for j in res["junctions"]]
    tmp, iso, perc, fam, cl, groups = op.cnt_percolation(lab, nl, bl, junc,
                                                         list(range(len(structure))))
    n = len(structure)
    a = [j[1] for j in junc]
    b = [j[3] for j in junc]
    _, comp = connected_components(coo_matrix((np.ones(len(a)), (a, b)), shape=(n, n)),
                                   directed=False)
    # clusters are the junction graph's components, one label each
    for L, tubes in enumerate(tmp):
        assert len({comp[t] for t in tubes}) == 1
    on = {b_: set(v) for b_, v in enumerate(bl)}
    for L, tubes in enumerate(tmp):
        ts = set(tubes)
        axes = {ax for ax, (lo, hi) in zip("xyz", ((4, 2), (5, 3), (1, 0)))
                if ts & on[lo] and ts & on[hi]}
        if axes:
            assert L in perc and fam[perc[L]] == FAMILY[frozenset(axes)]
        else:
            assert L not in perc and tubes in iso[1:]
    # a cluster on x_min (4) and y_max (3) only percolates in no direction; a restatement that
    # paired x_min with y_max (a swapped boundary number) would call it x-percolated
    t0 = tmp[0][0]
    only = [[], [], [], [t0], [t0], []]
    assert op.cnt_percolation(lab, nl, only, [], list(range(n)))[2] == {}
    swapped = [[], [], [t0], [], [t0], []]
    assert op.cnt_percolation(lab, nl, swapped, [], list(range(n)))[3] == [0]


def test_window_order_and_bad_input(case):
    xyz, structure, box, res, bl = case
    lab, nl = res["labels"], res["n_labels"]
    order = list(np.random.default_rng(3).permutation(len(structure)))
    got = clusters.cnt_percolation(lab, nl, bl, cnts_inside=order)
    tmp, iso, *_ = op.cnt_percolation(lab, nl, bl, [], order)
    assert [c.tolist() for c in got["clusters"]] == tmp and got["isolated"].tolist() == iso[0]
    # a junction across two clusters is an error, as in Group_junctions_same_particle
    t0 = next(t for t in range(len(lab)) if lab[t] == 0)
    t1 = next(t for t in range(len(lab)) if lab[t] == 1)
    with pytest.raises(RuntimeError, match="different clusters"):
        clusters.cnt_percolation(lab, nl, bl, np.array([[0, t0, 0, t1]]))
    with pytest.raises(RuntimeError, match="out of range"):
        clusters.cnt_percolation(lab, nl, [[len(lab)], [], [], [], [], []])
