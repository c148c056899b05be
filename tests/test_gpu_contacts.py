"""Tube contact search and cluster labels on the device (cicy_cnt_contacts) against
oracle.hk, exactly: labels and their count, every junction (points, tubes, distance bit for
bit, in order) and every element set -- on wavy-tube networks with the default background
grid, one subregion, a grid small enough to be resized (Adjust_regions_if_needed), tubes of
the window listed in a shuffled order, networks read back from the reference's file formats,
and the empty cases.  At a size the oracle does not reach, the labels are checked against
the connected components of the junction graph."""
import os

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import test_hk_oracle as th
from cicy_geom_amd import dea, netio

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dea.Context(0)
    yield c
    c.close()


def device(ctx, xyz, radii, structure, box, gs=0.05, minp=th.MINP, inside=None):
    g = th.grid_args(box, gs)
    return ctx.cnt_contacts(xyz, radii, structure, g["window_min"], g["window_len"],
                            g["gs_min"], g["overlap_cnt"], g["overlap_gnp"], th.VDW, th.TUNNEL,
                            minp, cnts_inside=inside)


def compare(got, want):
    assert got["labels"].tolist() == want["labels"]
    assert got["n_labels"] == want["n_labels"]
    assert len(got["junctions"]) == len(want["junctions"])
    assert [tuple(int(v) for v in j) for j in got["junctions"]] == \
        [w[:4] for w in want["junctions"]]
    assert got["junction_d"].tolist() == [w[4] for w in want["junctions"]]
    assert [e.tolist() for e in got["elements"]] == want["elements"]
    assert got["stats"]["contacts"] == len(want["contacts"])


CASES = {
    "default_grid": dict(seed=5, n=160, box=0.5, gs=0.05),
    "one_subregion": dict(seed=6, n=120, box=0.4, gs=0.4),
    "resized_grid": dict(seed=7, n=160, box=0.5, gs=0.01),  # 0.01 < 2 x overlap
    "dense": dict(seed=8, n=400, box=0.5, gs=0.05),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_matches_the_oracle(ctx, name):
    c = CASES[name]
    print("case", name, flush=True)
    xyz, radii, structure = th.wavy_tubes(c["seed"], c["n"], c["box"])
    got = device(ctx, xyz, radii, structure, c["box"], c["gs"])
    want = th.run(xyz, radii, structure, c["box"], gs=c["gs"])
    compare(got, want)
    assert want["n_labels"] > 1 and len(want["junctions"]) > 20


def test_shuffled_window_order_and_min_points(ctx):
    xyz, radii, structure = th.wavy_tubes(9, 200, 0.5)
    order = list(np.random.default_rng(2).permutation(len(structure))[:150])
    for minp in (1, 5, 10 ** 9):
        got = device(ctx, xyz, radii, structure, 0.5, minp=minp, inside=order)
        want = th.run(xyz, radii, structure, 0.5, minp=minp, inside=order)
        compare(got, want)


@pytest.mark.parametrize("fmt", ["csv", "dat"])
def test_network_read_from_files(ctx, tmp_path, fmt):
    """File -> reader -> contact search: the reference's input path for a stored network."""
    xyz, radii, structure = th.wavy_tubes(10, 160, 0.5)
    netio.write_network(tmp_path, fmt, [0, 0, 0], [0.5, 0.5, 0.5],
                        [len(s) for s in structure], radii, xyz, prec_cnt=6)
    corner, lengths = netio.read_sample_geometry(tmp_path)
    t = netio.read_cnts(tmp_path, fmt)
    got = device(ctx, t["xyz"], t["radii"], t["structure"], float(lengths[0]))
    want = th.run(t["xyz"], t["radii"], t["structure"], float(lengths[0]))
    compare(got, want)


def test_array_form_of_the_tubes(ctx):
    """(offsets, point ids) in, element arrays out: the same result as the list form."""
    xyz, radii, structure = th.wavy_tubes(12, 160, 0.5)
    off = np.concatenate([[0], np.cumsum([len(s) for s in structure])])
    g = th.grid_args(0.5)
    got = ctx.cnt_contacts(xyz, radii, (off, np.concatenate(structure)), g["window_min"],
                           g["window_len"], g["gs_min"], g["overlap_cnt"], g["overlap_gnp"],
                           th.VDW, th.TUNNEL, th.MINP, as_arrays=True)
    want = device(ctx, xyz, radii, structure, 0.5)
    assert np.array_equal(got["labels"], want["labels"])
    assert np.array_equal(got["junctions"], want["junctions"])
    eo, ep = got["elements_off"], got["elements_pts"]
    assert [ep[eo[i]:eo[i + 1]].tolist() for i in range(len(structure))] == \
        [e.tolist() for e in want["elements"]]


def test_empty_and_contactless(ctx):
    got = device(ctx, np.zeros((0, 3)), np.zeros(0), [], 0.5)
    assert got["n_labels"] == 0 and len(got["junctions"]) == 0
    xyz = np.array([[0.1, 0.1, 0.1], [0.2, 0.1, 0.1], [0.3, 0.3, 0.3], [0.4, 0.3, 0.3]])
    got = device(ctx, xyz, np.array([0.005, 0.005]), [np.arange(2), np.arange(2, 4)], 0.5)
    assert got["labels"].tolist() == [-1, -1] and got["n_labels"] == 0
    assert [e.tolist() for e in got["elements"]] == [[], []]


def test_bad_input_is_rejected(ctx):
    xyz = np.zeros((4, 3))
    with pytest.raises(dea.DEAError, match="two tubes"):
        device(ctx, xyz, np.array([0.005, 0.005]), [np.arange(3), np.arange(2, 4)], 0.5)
    with pytest.raises(dea.DEAError, match="out of range"):
        device(ctx, xyz, np.array([0.005]), [np.array([0, 7])], 0.5)
    with pytest.raises(dea.DEAError, match="invalid"):
        device(ctx, xyz, np.array([0.005]), [np.arange(4)], -1.0)
    bad = xyz.copy()
    bad[2, 1] = np.nan
    with pytest.raises(dea.DEAError, match="not finite"):
        device(ctx, bad, np.array([0.005]), [np.arange(4)], 0.5)


def test_large_network_matches_the_oracle(ctx):
    """0.59 M points: enough entries for the compression to run on several host threads (it
    splits above 65 536 entries); the result is still the oracle's, exactly."""
    xyz, radii, structure = th.wavy_tubes(11, 10_000, 2.0, n_min=40, n_max=120)
    got = device(ctx, xyz, radii, structure, 2.0)
    want = th.run(xyz, radii, structure, 2.0)
    assert len(want["contacts"]) > 3 * 65536
    compare(got, want)


def test_large_network_labels_follow_junctions(ctx):
    """2 um cube at ~6 % tubes (about 0.3 M points): tubes share a label exactly when the
    junctions connect them, labels run 0 .. n-1 in first-appearance order of their tubes'
    first contacts, and every junction joins two tubes of one label."""
    xyz, radii, structure = th.wavy_tubes(11, 10_000, 2.0, n_min=40, n_max=120)
    print("points", len(xyz), "tubes", len(structure), flush=True)
    got = device(ctx, xyz, radii, structure, 2.0)
    lab = got["labels"]
    j = got["junctions"]
    n = len(structure)
    g = coo_matrix((np.ones(len(j)), (j[:, 1], j[:, 3])), shape=(n, n))
    _, comp = connected_components(g, directed=False)
    touched = np.zeros(n, bool)
    touched[j[:, 1]] = touched[j[:, 3]] = True
    assert ((lab >= 0) == touched).all()
    assert (lab[j[:, 1]] == lab[j[:, 3]]).all()
    # same component <=> same label
    pairs = {(int(comp[i]), int(lab[i])) for i in np.nonzero(touched)[0]}
    assert len(pairs) == len({p[0] for p in pairs}) == len({p[1] for p in pairs})
    assert sorted({int(v) for v in lab if v >= 0}) == list(range(got["n_labels"]))
    assert (got["junction_d"] >= th.VDW).all() and (got["junction_d"] <= th.TUNNEL).all()
    print({k: round(v, 3) for k, v in got["stats"].items()}, "labels", got["n_labels"],
          "junctions", len(j), flush=True)


def test_percolation_from_device_labels(ctx):
    """Contact search on the device, then cicy_cnt_percolation: the same clusters, families
    and junction groups as the oracle chain (oracle.hk, then oracle.percolation)."""
    from cicy_geom_amd import clusters
    from oracle import percolation as op
    import test_percolation as tp
    xyz, radii, structure = th.wavy_tubes(21, 700, 0.5, n_min=30, n_max=90)
    bl = tp.boundary_lists(xyz, structure, 0.5, 0.011)
    got = device(ctx, xyz, radii, structure, 0.5)
    perc = clusters.cnt_percolation(got["labels"], got["n_labels"], bl, got["junctions"])
    want = th.run(xyz, radii, structure, 0.5)
    tmp, iso, pl, fam, cl, groups = op.cnt_percolation(
        want["labels"], want["n_labels"], bl, [j[:4] for j in want["junctions"]],
        list(range(len(structure))))
    assert perc["n_perc"] == len(pl) > 0
    assert [int(perc["family"][L]) for L in sorted(pl)] == fam
    assert [c.tolist() for c in perc["clusters"]] == tmp
    jc = perc["junction_cluster"]
    assert [list(np.nonzero(jc == k)[0]) for k in range(len(groups))] == groups
