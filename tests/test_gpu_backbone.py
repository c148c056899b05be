"""Backbone scan on the device (cicy_backbone_scan): the zero current and every tube's backbone
run match oracle.backbone.scan exactly -- on clusters with tubes only and with plates of both
sizes, on ragged synthetic input (empty and one-point tubes, no other elements, no tubes), and
end to end: unit-resistor solve on the device, scan on the device, host pruning with the
device's zero current and runs, against the oracle's extraction on the same voltages."""
import numpy as np
import pytest

import test_backbone as tb
from oracle import backbone as ob
from cicy_geom_amd import dea, network

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dea.Context(0)
    yield c
    c.close()


def node_lists(off, nodes):
    return [[int(x) for x in nodes[off[t]:off[t + 1]]] for t in range(len(off) - 1)]


def check_scan(ctx, v, off, nodes, pa, pb):
    zero, seg = ctx.backbone_scan(v, off, nodes, pa, pb)
    want_zero, want_runs = ob.scan(v, node_lists(off, nodes),
                                   list(zip(pa.tolist(), pb.tolist())))
    assert zero == want_zero
    assert [tuple(int(x) for x in r) for r in seg] == want_runs
    return zero, seg


@pytest.mark.parametrize("name", sorted(tb.CASES))
def test_scan_matches_the_oracle_on_clusters(ctx, name):
    print("case", name, flush=True)
    net = network.Network(tb.CASES[name])
    k = tb.largest(net)
    s0 = net.assemble(k, 0, net.cluster(k)[1])
    v = tb.solve(s0)
    off, nodes, pa, pb = net.backbone_scan_inputs(k, s0)
    zero, seg = check_scan(ctx, v, off, nodes, pa, pb)
    assert (seg[:, 0] == -1).any() and (seg[:, 0] >= 0).any()


def test_scan_ragged_synthetic(ctx):
    rng = np.random.default_rng(7)
    n = 50_000
    v = rng.standard_normal(n)
    # flat stretches: runs of equal voltages make dead elements inside and at tube ends
    flat = rng.integers(0, n - 40, 400)
    for f in flat:
        v[f:f + rng.integers(2, 40)] = v[f]
    lens = rng.integers(0, 60, 3000)
    lens[:5] = [0, 1, 0, 2, 1]
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    start = rng.integers(0, n - 60, len(lens))
    nodes = np.concatenate([np.arange(s, s + l) for s, l in zip(start, lens)]).astype(np.int64)
    pa = rng.integers(0, n, 20_000)
    pb = rng.integers(0, n, 20_000)
    check_scan(ctx, v, off, nodes, pa, pb)
    # no other elements; no tubes; a maximum that only a pair carries
    check_scan(ctx, v, off, nodes, pa[:0], pb[:0])
    check_scan(ctx, v, np.zeros(1, np.int64), nodes[:0], pa, pb)
    assert pa[0] != pb[0]
    w = np.zeros(n)
    w[pa[0]], w[pb[0]] = 1.0, -1.0
    zero, _ = check_scan(ctx, w, off, nodes, pa[:1], pb[:1])
    assert zero == 2e-9
    # no current anywhere: the zero current is 1e-9 times the floor of 1e-10, nothing lives
    zero, seg = check_scan(ctx, np.zeros(n), off, nodes, pa, pb)
    assert zero == 1e-10 * 1e-9 and (seg == -1).all()  # the product, not 1e-19


def test_scan_rejects_bad_input(ctx):
    v = np.zeros(10)
    off = np.array([0, 3], np.int64)
    with pytest.raises(dea.DEAError):
        ctx.backbone_scan(v, off, np.array([0, 1, 10], np.int64), [], [])
    with pytest.raises(dea.DEAError):
        ctx.backbone_scan(v, off, np.array([0, 1, 2], np.int64), [0], [-1])
    with pytest.raises(dea.DEAError):
        ctx.backbone_scan(v, np.array([0, 2, 1], np.int64), np.arange(2), [], [])


@pytest.mark.parametrize("name", ["tubes", "plates"])
def test_backbone_end_to_end_on_the_device(ctx, name):
    """Solve the unit-resistor system with the drop-in, scan on the device, prune on the host
    with the device's zero current and runs: the result equals the oracle's extraction from
    the same voltages."""
    print("case", name, flush=True)
    net = network.Network(tb.CASES[name])
    k = tb.largest(net)
    s0 = net.assemble(k, 0, net.cluster(k)[1])
    ok, volt, x, R, st = ctx.solve_dea_cg_sss(
        s0.nodes, s0.reserved, 1e-10, s0.col_ind, s0.row_ptr, s0.values, s0.diag, s0.r0,
        s0.vef, s0.x0, max_iter=100 * s0.n)
    assert ok
    before = tb.snapshot(net, k)
    want_vol, want_ed = ob.backbone(before, volt, s0.lmm, s0.lmm_gnp, s0.reserved)
    off, nodes, pa, pb = net.backbone_scan_inputs(k, s0)
    zero, seg = check_scan(ctx, volt, off, nodes, pa, pb)
    got = net.backbone(k, s0, volt, zero=zero, segments=seg)
    assert got == want_vol
    assert [int(t) for t in net.cluster_tubes(k)] == want_ed["cluster"]
    for t, pts in want_ed["elements"].items():
        assert [int(p) for p in net.elements(t)] == pts, t
