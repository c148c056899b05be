"""Backbone extraction (Backbone_Network.cpp) on the host: cicy_net_backbone against the
restatement in oracle/backbone.py, and three checks that neither restatement can satisfy by
repeating the other's mistake:

* the tube volumes split exactly: backbone + dead = every cluster tube from its first to
  its last element point;
* every element the backbone keeps carries more than the zero current, and every tube it
  drops carries none above it;
* trimming dead branches leaves the cluster's resistance unchanged (they carry no current).

No GPU: the unit-resistor voltages come from a sparse direct solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from oracle import backbone as ob
from cicy_geom_amd import network


def solve(s):
    n = s.n
    rows = np.repeat(np.arange(n), np.diff(s.row_ptr))
    L = sp.csr_matrix((s.values, (rows, s.col_ind)), shape=(n, n))
    A = (L + L.T + sp.diags(s.diag)).tocsc()
    return np.concatenate([s.vef, sla.spsolve(A, s.b)])


def resistance(net, k, direction):
    s = net.assemble(k, 1, direction if direction < 3 else 0, direction=direction)
    v = solve(s)
    i1 = oracle.boundary_current(*s.boundary_elements[0], v)
    return abs(1.0 / i1)


def snapshot(net, k):
    """The cluster's data as oracle.backbone reads it."""
    tubes = [int(t) for t in net.cluster_tubes(k)]
    gnps = [int(g) for g in net.cluster_gnps(k)]
    junc = [net.junction(int(j))[:4] for j in net.cluster_junctions(k)]
    mixed = [net.mixed_junction(int(j))[:4] for j in net.cluster_mixed_junctions(k)]
    gj = [net.gnp_junction(int(j))[:4] for j in net.cluster_gnp_junctions(k)]
    touched = set(tubes) | {m[2] for m in mixed} | {j[2] for j in junc} | {j[3] for j in junc}
    plates = {}
    for g in gnps:
        c, ax, l, t = net.gnp(g)
        plates[g] = dict(centre=list(c), axes=[list(r) for r in ax], side=l, thick=t)
    return dict(xyz=net.xyz.tolist(), radius={t: float(net.radii[t]) for t in touched},
                box=net.params.box,
                elements={t: [int(p) for p in net.elements(t)] for t in touched},
                cluster=tubes, junctions=[tuple(int(x) for x in j) for j in junc],
                gnps=gnps, structure={g: [int(p) for p in net.gnp_structure(g)] for g in gnps},
                triangulation={g: [tuple(int(x) for x in e) for e in net.gnp_triangulation(g)]
                               for g in gnps},
                gjunctions=[tuple(int(x) for x in j) for j in gj],
                mixed=[tuple(int(x) for x in m) for m in mixed], plates=plates)


def largest(net):
    return max(range(net.count("clusters")), key=lambda c: net.cluster(c)[0])


def extract(params, avoid=False):
    """(net, cluster, snapshot before, oracle result, product result, unit voltages)"""
    net = network.Network(params)
    k = largest(net)
    fam = net.cluster(k)[1]
    s0 = net.assemble(k, 0, fam)
    v = solve(s0)
    before = snapshot(net, k)
    want = ob.backbone(before, v, s0.lmm, s0.lmm_gnp, s0.reserved, avoid_resistance=avoid)
    got = net.backbone(k, s0, v, avoid_resistance=avoid)
    return net, k, before, want, got, (s0, v)


CASES = {
    "tubes": network.NetworkParams(box=1.0, n_cnt=300, seed=1_000_000),
    # 0.3 um plates: some hang off a single tube and die, with junctions of both kinds
    "plates": network.NetworkParams(box=1.5, n_cnt=1080, n_gnp=150, gnp_side=0.3, seed=24),
    "large_plates": network.NetworkParams(box=1.5, n_cnt=1080, n_gnp=14, seed=11),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    return request.param, extract(CASES[request.param])


def test_matches_the_restatement(case):
    name, (net, k, before, want, got, _) = case
    vol, ed = want
    for key in vol:
        assert got[key] == vol[key], (key, got[key], vol[key])
    assert [int(t) for t in net.cluster_tubes(k)] == ed["cluster"]
    for t, pts in ed["elements"].items():
        assert [int(p) for p in net.elements(t)] == pts, t
    assert [net.junction(int(j))[:4] for j in net.cluster_junctions(k)] == ed["junctions"]
    assert [net.mixed_junction(int(j))[:4] for j in net.cluster_mixed_junctions(k)] == \
        ed["mixed"]
    assert [net.gnp_junction(int(j))[:4] for j in net.cluster_gnp_junctions(k)] == \
        ed["gjunctions"]
    for g, pts in ed["structure"].items():
        assert [int(p) for p in net.gnp_structure(g)] == pts, g
    print(name, {k_: round(v_, 9) for k_, v_ in got.items()},
          len(before["cluster"]), "->", len(ed["cluster"]), "tubes")
    assert len(ed["cluster"]) < len(before["cluster"])  # these clusters have dead ends
    if name == "plates":
        # every pruning path ran: a dead plate, mixed and plate-plate junctions removed
        assert got["gnp_backbone"] > 0.0 and got["gnp_dead"] > 0.0
        assert len(ed["mixed"]) < len(before["mixed"])
        assert len(ed["gjunctions"]) < len(before["gjunctions"])


def test_tube_volumes_split_exactly(case):
    name, (net, k, before, want, got, _) = case
    _, ed = want
    xyz = before["xyz"]
    total = dead1 = dead2 = run = 0.0
    for t in reversed(before["cluster"]):
        e, r = before["elements"][t], before["radius"][t]
        total += ob.cnt_volume(xyz, e[0], e[-1], r)
        i1, i2 = ed["segments"][t]
        if i1 == -1:
            dead1 += ob.cnt_volume(xyz, e[0], e[-1], r)
            continue
        dead1 += ob.cnt_volume(xyz, e[0], e[i1], r)
        run += ob.cnt_volume(xyz, e[i1], e[i2] if i2 >= 0 else e[-1], r)
        if i2 >= 0:
            dead2 += ob.cnt_volume(xyz, e[i2], e[-1], r)
    assert got["cnt_backbone"] + got["cnt_dead"] == pytest.approx(total, rel=1e-13)
    assert got["cnt_backbone"] == pytest.approx(run, rel=1e-13)
    # a planted mistake: a restatement that dropped the dead end after the backbone run
    # would report dead1 only, and the split would no longer add up
    assert dead2 > 0.0
    assert abs(got["cnt_backbone"] + dead1 - total) > 1e-6 * total


def test_kept_elements_carry_current(case):
    name, (net, k, before, want, got, (s0, v)) = case
    zero = got["zero_current"]
    lmm = s0.lmm
    kept = set(int(t) for t in net.cluster_tubes(k))
    for t in before["cluster"]:
        e = [int(p) for p in net.elements(t)]
        cur = [abs(v[lmm[e[q - 1]]] - v[lmm[e[q]]]) for q in range(1, len(e))]
        if t in kept:
            assert len(e) >= 2
            # the run ends at its first dead element, which is kept as the last point
            assert all(c - zero > ob.ZERO for c in cur[:-1]), t
        else:
            e0 = before["elements"][t]
            cur0 = [abs(v[lmm[e0[q - 1]]] - v[lmm[e0[q]]]) for q in range(1, len(e0))]
            touches_plate = any(m[2] == t for m in before["mixed"])
            assert touches_plate or all(c - zero <= ob.ZERO for c in cur0), t


@pytest.mark.parametrize("name,rel", [("tubes", 1e-9), ("large_plates", 1e-5),
                                      ("plates", 3e-3)])
def test_resistance_unchanged_by_trimming(name, rel):
    """Dead branches carry no current: the backbone's resistance equals the cluster's.  Where
    plates lose junction points their triangulation is made again from fewer points, as the
    reference's next assembly does (Direct_Electrifying.cpp:1037-1040), which changes their
    internal conductance a little: there the bar is 1e-5 (few points lost) or 0.3 %."""
    net = network.Network(CASES[name])
    k = largest(net)
    n_full = net.assemble(k, 1, 0, direction=0).n
    r_full = resistance(net, k, 0)
    fam = net.cluster(k)[1]
    s0 = net.assemble(k, 0, fam)
    net.backbone(k, s0, solve(s0))
    r_bb = resistance(net, k, 0)
    n_bb = net.assemble(k, 1, 0, direction=0).n
    print(name, "R full", r_full, "backbone", r_bb, "nodes", n_full, "->", n_bb)
    assert n_bb < n_full
    assert abs(r_bb - r_full) <= rel * r_full


def test_avoid_resistance_keeps_the_cluster():
    net, k, before, want, got, _ = extract(CASES["tubes"], avoid=True)
    vol, ed = want
    assert got == pytest.approx(vol, rel=0, abs=0)
    assert [int(t) for t in net.cluster_tubes(k)] == before["cluster"]
    for t in before["cluster"]:
        assert [int(p) for p in net.elements(t)] == before["elements"][t]


def test_bad_input_is_rejected():
    net = network.Network(CASES["tubes"])
    k = largest(net)
    s0 = net.assemble(k, 0, net.cluster(k)[1])
    with pytest.raises(RuntimeError):
        net.backbone(k, s0, np.zeros(3))  # voltages too short for the nodes


def test_scan_agrees_with_the_extraction(case):
    """oracle.backbone.scan (the device scan's checker) gives the zero current and the tube
    runs that the full extraction found, from the cluster's node lists and voltages."""
    name, (net, k, before, want, got, (s0, v)) = case
    vol, ed = want
    lmm, lg = s0.lmm, s0.lmm_gnp
    tubes = [[int(lmm[p]) for p in before["elements"][t]] for t in before["cluster"]]
    pairs = [(lmm[a], lmm[b]) for a, b, _, _ in before["junctions"]]
    pairs += [(lg[a], lg[b]) for g in before["gnps"] for a, b in before["triangulation"][g]]
    pairs += [(lg[a], lg[b]) for a, b, _, _ in before["gjunctions"]]
    pairs += [(lmm[a], lg[b]) for a, b, _, _ in before["mixed"]]
    zero, runs = ob.scan(v, tubes, pairs)
    assert zero == vol["zero_current"]
    assert runs == [ed["segments"][t] for t in before["cluster"]]
    # the helper that feeds the device scan lists the same nodes
    off, nodes, pa, pb = net.backbone_scan_inputs(k, s0)  # after the edit: fewer tubes
    assert off[-1] == len(nodes) and len(pa) == len(pb)
