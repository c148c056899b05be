"""oracle.hk (restatement of Contact_grid.cpp's tube subregions and Hoshen_Kopelman.cpp's
Make_cnt_clusters) pinned by checks it cannot pass by repeating its own mistakes:

* with one subregion the contacts are exactly the brute-force pairs (a k-d tree search)
  that pass the tunnelling rule; with many, every brute-force pair whose points share a home
  subregion is found and nothing else is;
* the labels are the connected components of the contact graph, numbered in the order of
  each component's first contact (what HK76 with smaller-root merging and Cleanup_labels
  produce), computed here with scipy's connected_components;
* the compression rule at its two ends: min_points = 1 gives one junction per (point, other
  tube), an unreachable min_points one per tube pair, each at the pair's closest points;
* two crossing straight tubes give one junction at the known closest points.
The networks are network.wavy_tubes (0.01 um steps, the reference's default)."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial import cKDTree

from oracle import hk
from cicy_geom_amd import network

VDW, TUNNEL, MINP = 0.00034, 0.0018, 5  # input.dat defaults (Cutoff_Distances)


wavy_tubes = network.wavy_tubes


def grid_args(box, gs=0.05, rmax=0.006):
    return dict(window_min=[0.0, 0.0, 0.0], window_len=[box] * 3, gs_min=[gs] * 3,
                overlap_cnt=2 * rmax + TUNNEL, overlap_gnp=0.0)


def run(xyz, radii, structure, box, gs=0.05, minp=MINP, inside=None):
    g = grid_args(box, gs)
    return hk.cnt_clusters(g["window_min"], g["window_len"], g["gs_min"], g["overlap_cnt"],
                           g["overlap_gnp"], (VDW, TUNNEL, minp), radii, structure, xyz,
                           cnts_inside=inside)


def brute_pairs(xyz, radii, structure):
    flag = np.concatenate([np.full(len(s), i) for i, s in enumerate(structure)])
    tree = cKDTree(xyz)
    out = set()
    for a, b in tree.query_pairs(2 * radii.max() + TUNNEL + 1e-9):
        if flag[a] == flag[b]:
            continue
        d = np.sqrt(((xyz[a] - xyz[b]) ** 2).sum()) - radii[flag[a]] - radii[flag[b]]
        if max(d, VDW) <= TUNNEL:
            out.add((min(a, b), max(a, b)))
    return out, flag


@pytest.fixture(scope="module")
def net():
    return wavy_tubes(5, 160, 0.5)


def test_one_subregion_finds_every_pair(net):
    xyz, radii, structure = net
    res = run(xyz, radii, structure, 0.5, gs=0.5)
    assert len(res["regions"]) == 1
    found = {(min(c[3], c[4]), max(c[3], c[4])) for c in res["contacts"]}
    want, _ = brute_pairs(xyz, radii, structure)
    assert len(want) > 100 and found == want


def test_subregions_find_pairs_that_share_a_home(net):
    xyz, radii, structure = net
    res = run(xyz, radii, structure, 0.5)
    assert len(res["regions"]) == 1000
    found = {(min(c[3], c[4]), max(c[3], c[4])) for c in res["contacts"]}
    want, _ = brute_pairs(xyz, radii, structure)
    assert found <= want
    home = np.minimum((xyz / 0.05).astype(int), 9)
    key = home[:, 0] + 10 * home[:, 1] + 100 * home[:, 2]
    same = {p for p in want if key[p[0]] == key[p[1]]}
    assert same <= found and len(same) > 50
    # a pair split across a face is found through the overlap; most are
    assert len(found) > 0.9 * len(want)


def labels_by_components(contacts, flag, n_cnts):
    a = np.array([flag[c[3]] for c in contacts])
    b = np.array([flag[c[4]] for c in contacts])
    g = coo_matrix((np.ones(len(a)), (a, b)), shape=(n_cnts, n_cnts))
    _, comp = connected_components(g, directed=False)
    first = {}
    for q, (x, y) in enumerate(zip(a, b)):
        first.setdefault(comp[x], q)
    rank = {c: r for r, c in enumerate(sorted(first, key=first.get))}
    touched = set(a) | set(b)
    return [rank[comp[i]] if i in touched else -1 for i in range(n_cnts)], len(rank)


@pytest.mark.parametrize("gs,inside", [(0.05, None), (0.5, None), (0.05, "shuffled")])
def test_labels_are_components_in_first_contact_order(net, gs, inside):
    xyz, radii, structure = net
    order = None
    if inside == "shuffled":
        order = list(np.random.default_rng(1).permutation(len(structure)))
    res = run(xyz, radii, structure, 0.5, gs=gs, inside=order)
    _, flag = brute_pairs(xyz, radii, structure)
    want, n = labels_by_components(res["contacts"], flag, len(structure))
    assert res["labels"] == want and res["n_labels"] == n
    assert n > 3


def test_compression_rule_at_its_ends(net):
    xyz, radii, structure = net
    _, flag = brute_pairs(xyz, radii, structure)
    every = run(xyz, radii, structure, 0.5, minp=1)
    keys = set()
    for c in every["contacts"]:
        P1, P2 = c[3], c[4]
        a, b = (P1, flag[P2]) if flag[P1] < flag[P2] else (P2, flag[P1])
        keys.add((a, b))
    assert len(every["junctions"]) == len(keys)
    one = run(xyz, radii, structure, 0.5, minp=10 ** 9)
    best = {}
    for c in one["contacts"]:
        i, j = sorted((flag[c[3]], flag[c[4]]))
        if (i, j) not in best or c[5] < best[(i, j)]:
            best[(i, j)] = c[5]
    assert len(one["junctions"]) == len(best)
    for Pi, i, Pj, j, d in one["junctions"]:
        assert i < j and flag[Pi] == i and flag[Pj] == j
        assert d == best[(i, j)]
    # the default rule lies between the two
    mid = run(xyz, radii, structure, 0.5)
    assert len(best) <= len(mid["junctions"]) <= len(keys)
    for Pi, i, Pj, j, d in mid["junctions"]:
        assert Pi in mid["elements"][i] and Pj in mid["elements"][j]
    for i, e in enumerate(mid["elements"]):
        if e:
            assert e[0] == structure[i][0] and e[-1] == structure[i][-1]


def test_two_crossing_tubes():
    """Tube 0 along x at z = 0, tube 1 along y at z = 0.0115: radii 0.005, so the gap at the
    crossing is 0.0015 <= 0.0018 and exactly one point pair (the middle ones) is in contact."""
    s = np.linspace(-0.05, 0.05, 11) + 0.25
    t0 = np.stack([s, np.full(11, 0.25), np.full(11, 0.25)], 1)
    t1 = np.stack([np.full(11, 0.25), s, np.full(11, 0.2615)], 1)
    xyz = np.concatenate([t0, t1])
    structure = [np.arange(11), np.arange(11, 22)]
    radii = np.array([0.005, 0.005])
    res = run(xyz, radii, structure, 0.5)
    assert res["labels"] == [0, 0] and res["n_labels"] == 1
    assert len(res["junctions"]) == 1
    Pi, i, Pj, j, d = res["junctions"][0]
    assert (Pi, i, Pj, j) == (5, 0, 16, 1)
    assert d == pytest.approx(0.0115 - 0.01, abs=1e-15)
    assert res["elements"] == [[0, 5, 10], [11, 16, 21]]
