"""Host forms of the condensed solve and of the resident layout, checked against dense
linear algebra and closed forms.  No GPU calls."""
import numpy as np
import pytest

import oracle
from cicy_geom_amd import dea, network


def dense(rp, col, val, diag):
    n = len(diag)
    A = np.diag(np.asarray(diag, float))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            A[r, col[k]] += val[k]
            A[col[k], r] += val[k]
    return A


@pytest.fixture(scope="module")
def c1():
    return network.largest_percolating_system(network.generate(1), 0)


def test_reduced_matrix_is_the_schur_complement(c1):
    s = c1
    anchors, rrp, rcol, rval, rdiag, rrhs = dea.host_condense(s.row_ptr, s.col_ind, s.values,
                                                              s.diag, s.b)
    A = dense(s.row_ptr, s.col_ind, s.values, s.diag)
    n = s.n
    J = anchors
    I = np.setdiff1d(np.arange(n), J)
    # the anchors are exactly the rows with a junction (off-sub-diagonal) entry
    rows = np.repeat(np.arange(n), np.diff(s.row_ptr))
    off = s.col_ind != rows - 1
    assert set(J) == set(rows[off]) | set(s.col_ind[off])
    AII, AIJ, AJJ = A[np.ix_(I, I)], A[np.ix_(I, J)], A[np.ix_(J, J)]
    S = AJJ - AIJ.T @ np.linalg.solve(AII, AIJ)
    Sred = dense(rrp, rcol, rval, rdiag)
    assert np.abs(Sred - S).max() <= 1e-10 * np.abs(S).max()
    bred = s.b[J] - AIJ.T @ np.linalg.solve(AII, s.b[I])
    assert np.allclose(rrhs, bred, rtol=1e-10, atol=1e-12 * np.abs(bred).max())
    # reduced matrix stays strictly lower SSS with the anchors in order
    assert np.all(np.diff(J) > 0)
    for r in range(len(J)):
        assert np.all(rcol[rrp[r]:rrp[r + 1]] < r)


def test_expand_recovers_the_full_solution(c1):
    s = c1
    anchors, rrp, rcol, rval, rdiag, rrhs = dea.host_condense(s.row_ptr, s.col_ind, s.values,
                                                              s.diag, s.b)
    xj = np.linalg.solve(dense(rrp, rcol, rval, rdiag), rrhs)
    x = dea.host_expand(s.row_ptr, s.col_ind, s.values, s.diag, s.b, xj)
    want = np.linalg.solve(dense(s.row_ptr, s.col_ind, s.values, s.diag), s.b)
    assert np.linalg.norm(x - want) <= 1e-10 * np.linalg.norm(want)
    ax = oracle.spmv_sss(s.row_ptr, s.col_ind, s.values, s.diag, x)
    assert np.linalg.norm(s.b - ax) <= 1e-12 * np.linalg.norm(s.b)


def test_core_is_the_schur_complement_and_expands_exactly(c1):
    """Second stage (series rounds): the core matrix is the Schur complement of A onto the
    core rows, its right-hand side the condensed one, and the back-substitution of both
    stages gives the full solution."""
    s = c1
    rows, crp, ccol, cval, cdiag, crhs, rounds = dea.host_core(s.row_ptr, s.col_ind, s.values,
                                                               s.diag, s.b)
    anchors = dea.host_condense(s.row_ptr, s.col_ind, s.values, s.diag)[0]
    assert rounds >= 1 and 0 < len(rows) < len(anchors)
    assert set(rows) <= set(anchors) and np.all(np.diff(rows) > 0)
    A = dense(s.row_ptr, s.col_ind, s.values, s.diag)
    J = rows
    I = np.setdiff1d(np.arange(s.n), J)
    AII, AIJ, AJJ = A[np.ix_(I, I)], A[np.ix_(I, J)], A[np.ix_(J, J)]
    S = AJJ - AIJ.T @ np.linalg.solve(AII, AIJ)
    C = dense(crp, ccol, cval, cdiag)
    assert np.abs(C - S).max() <= 1e-10 * np.abs(S).max()
    bc = s.b[J] - AIJ.T @ np.linalg.solve(AII, s.b[I])
    assert np.allclose(crhs, bc, rtol=1e-9, atol=1e-12 * np.abs(bc).max())
    xc = np.linalg.solve(C, crhs)
    x = dea.host_core_expand(s.row_ptr, s.col_ind, s.values, s.diag, s.b, xc)
    want = np.linalg.solve(A, s.b)
    assert np.linalg.norm(x - want) <= 1e-10 * np.linalg.norm(want)


def test_star_reduces_completely():
    """A hub with three spokes of series resistors to ground is series-parallel: both stages
    eliminate every row, and the back-substitution alone solves it exactly (hub voltage of
    a unit current source = 1 / (3 g/3))."""
    g = 2.0
    n = 7          # hub 0; spoke s: node 1+2s next to the hub, node 2+2s grounded through g
    lower = {}
    for k in range(3):
        a, b = 1 + 2 * k, 2 + 2 * k
        lower[(a, 0)] = -g
        lower[(b, a)] = -g
    keys = sorted(lower)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int64)
    val = np.array([lower[k] for k in keys])
    diag = np.array([3 * g] + [2 * g, 2 * g] * 3)
    rhs = np.zeros(n)
    rhs[0] = 1.0
    rows, *_ = dea.host_core(rp, col, val, diag, rhs)
    assert len(rows) == 0
    x = dea.host_core_expand(rp, col, val, diag, rhs, np.zeros(0))
    assert x[0] == pytest.approx(1.0 / (3 * g / 3.0), rel=1e-13)
    assert np.allclose(x, np.linalg.solve(dense(rp, col, val, diag), rhs), rtol=1e-13)


def chain_system(k, g, junction=0.5):
    """Two anchors (rows 0 and k+1) joined by a chain of k interior rows of conductance g,
    plus a junction of conductance `junction` between the anchors and a ground leak on
    each anchor so the system is definite."""
    n = k + 2
    rp, col, val = [0], [], []
    for r in range(n):
        if r == k + 1:
            col.append(0)
            val.append(-junction)
        if r >= 1:
            col.append(r - 1)
            val.append(-g)
        rp.append(len(col))
    # sort each row's columns
    rp = np.array(rp, np.int64)
    col = np.array(col, np.int64)
    val = np.array(val)
    for r in range(n):
        o = np.argsort(col[rp[r]:rp[r + 1]])
        col[rp[r]:rp[r + 1]] = col[rp[r]:rp[r + 1]][o]
        val[rp[r]:rp[r + 1]] = val[rp[r]:rp[r + 1]][o]
    diag = np.full(n, 2 * g)
    diag[0] = diag[-1] = g + junction + 1.0
    return rp, col, val, diag


@pytest.mark.parametrize("k", [2, 3, 7, 40])
def test_series_law(k):
    """k equal conductances g in series between two anchors act as one of g/(k+1)."""
    g = 3.0
    rp, col, val, diag = chain_system(k, g)
    anchors, rrp, rcol, rval, rdiag, _ = dea.host_condense(rp, col, val, diag)
    assert list(anchors) == [0, k + 1]
    geff = g / (k + 1)
    assert rval[0] == pytest.approx(-0.5 - geff, rel=1e-13)   # junction + series path
    # each anchor keeps its leak and junction, and sees the path as geff to the other
    assert rdiag == pytest.approx([1.0 + 0.5 + geff] * 2, rel=1e-13)


def floating_system(k=6, g=3.0, m=5):
    """A definite chain (rows 0 .. k+1) plus a floating chain of m rows that touches no
    junction and no electrode (singular on its own)."""
    rp0, col0, val0, diag0 = chain_system(k, g)
    n0 = len(diag0)
    rp = list(rp0)
    col, val = list(col0), list(val0)
    for r in range(m):
        if r > 0:
            col.append(n0 + r - 1)
            val.append(-g)
        rp.append(len(col))
    rp, col, val = np.array(rp, np.int64), np.array(col, np.int64), np.array(val)
    diag = np.concatenate([diag0, [g] + [2 * g] * (m - 2) + [g]])
    return (rp0, col0, val0, diag0), (rp, col, val, diag), n0


def test_floating_tube_from_a_ramp_ends_at_its_mean():
    """From an initial guess x0 that varies along a floating tube, r0 = -A x0 there (the
    block is singular but consistent).  CG's correction stays orthogonal to the constant
    null vector, so the tube ends at mean(x0); the condensed correction must be exactly
    mean(x0) - x0 there, and it must not disturb the definite part."""
    (rp0, col0, val0, diag0), (rp, col, val, diag), n0 = floating_system()
    m = len(diag) - n0
    x0 = np.zeros(n0 + m)
    x0[n0:] = np.linspace(0.9, 0.1, m)             # a ramp over the floating tube
    b = np.zeros(n0 + m)
    b[0] = 1.0
    A = dense(rp, col, val, diag)
    r0 = b - A @ x0
    rows, crp, ccol, cval, cdiag, crhs, _ = dea.host_core(rp, col, val, diag, r0)
    xc = np.linalg.solve(dense(crp, ccol, cval, cdiag), crhs) if len(rows) else np.zeros(0)
    d = dea.host_core_expand(rp, col, val, diag, r0, xc)
    assert np.allclose(d[n0:], x0[n0:].mean() - x0[n0:], rtol=0, atol=1e-14)
    x = x0 + d
    assert np.linalg.norm(b - A @ x) <= 1e-13
    want = np.linalg.solve(dense(rp0, col0, val0, diag0), b[:n0])
    assert np.allclose(x[:n0], want, rtol=1e-12)
    # planted mistake: leaving the tube at x0 (no correction) keeps the residual -A x0
    assert np.linalg.norm(b - A @ (x0 + np.concatenate([d[:n0], np.zeros(m)]))) > 1e-3


def test_floating_tube_is_left_at_zero():
    """A tube touching no junction and no electrode is singular on its own; CG from zero
    leaves it at zero and so does the condensation (the rest is solved exactly)."""
    k, g = 6, 3.0
    rp0, col0, val0, diag0 = chain_system(k, g)          # rows 0 .. k+1, definite
    n0 = len(diag0)
    m = 5                                                # floating chain rows n0 .. n0+m-1
    rp = list(rp0)
    col, val = list(col0), list(val0)
    for r in range(m):
        if r > 0:
            col.append(n0 + r - 1)
            val.append(-g)
        rp.append(len(col))
    rp, col, val = np.array(rp, np.int64), np.array(col, np.int64), np.array(val)
    diag = np.concatenate([diag0, [g] + [2 * g] * (m - 2) + [g]])
    rhs = np.zeros(n0 + m)
    rhs[0] = 1.0
    rows, crp, ccol, cval, cdiag, crhs, _ = dea.host_core(rp, col, val, diag, rhs)
    xc = np.linalg.solve(dense(crp, ccol, cval, cdiag), crhs) if len(rows) else np.zeros(0)
    x = dea.host_core_expand(rp, col, val, diag, rhs, xc)
    assert np.all(x[n0:] == 0.0)
    want = np.linalg.solve(dense(rp0, col0, val0, diag0), rhs[:n0])
    assert np.allclose(x[:n0], want, rtol=1e-12)


def test_nothing_to_condense_is_refused():
    rng = np.random.default_rng(3)
    n = 200
    pairs = {}
    for a, b in rng.integers(0, n, (800, 2)):
        if abs(a - b) > 1:
            pairs[(max(a, b), min(a, b))] = -1.0
    keys = sorted(pairs)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int64)
    with pytest.raises(dea.DEAError):
        dea.host_condense(rp, col, -np.ones(len(col)), np.full(n, 10.0))


def band_system(n, seed):
    """Tube-like system: chains along the sub-diagonal with breaks, random junctions."""
    rng = np.random.default_rng(seed)
    lower = {}
    for r in range(1, n):
        if rng.random() > 0.05:
            lower[(r, r - 1)] = -rng.uniform(0.5, 2.0)
    for _ in range(n // 12):
        a, b = rng.integers(0, n, 2)
        if abs(a - b) > 1:
            lower[(max(a, b), min(a, b))] = -10.0 ** rng.uniform(-6, 0)
    keys = sorted(lower)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int64)
    val = np.array([lower[k] for k in keys])
    diag = np.full(n, 0.1)
    for (r, c), v in lower.items():
        diag[r] -= v
        diag[c] -= v
    return rp, col, val, diag


@pytest.mark.parametrize("R", [4, 8])
@pytest.mark.parametrize("phase", [0, 1])
def test_resident_layout_on_host(R, phase):
    """The resident kernel's walk (lanes, wave edges, published rows of neighbouring
    work-groups, remainder slots) gives A^ (r + beta p_old) over several work-groups."""
    n = 512 * R * 3 + 777           # four work-groups, the last one ragged
    rp, col, val, diag = band_system(n, R + phase)
    rng = np.random.default_rng(7)
    r = rng.standard_normal(n)
    pold = rng.standard_normal(n)
    beta = 0.37
    got = dea.host_resident_apply(rp, col, val, diag, r, pold, beta, R, phase)
    sc = 1.0 / np.sqrt(diag)
    p = r + beta * pold
    want = sc * oracle.spmv_sss(rp, col, val, diag, sc * p)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("phase", [0, 1])
def test_resident_layout_without_band_structure(phase):
    """A junction-dominated matrix (every row published), as a core system is."""
    rng = np.random.default_rng(11 + phase)
    n = 2048 * 2 + 300
    lower = {}
    for a, b in rng.integers(0, n, (3 * n, 2)):
        if a != b:
            lower[(max(a, b), min(a, b))] = -10.0 ** rng.uniform(-4, 0)
    keys = sorted(lower)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int64)
    val = np.array([lower[k] for k in keys])
    diag = np.full(n, 0.01)
    for (r, c), v in lower.items():
        diag[r] -= v
        diag[c] -= v
    r = rng.standard_normal(n)
    pold = rng.standard_normal(n)
    got = dea.host_resident_apply(rp, col, val, diag, r, pold, -0.8, 4, phase)
    sc = 1.0 / np.sqrt(diag)
    want = sc * oracle.spmv_sss(rp, col, val, diag, sc * (r - 0.8 * pold))
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_resident_layout_rejects_unknown_rows_per_thread(c1):
    s = c1
    with pytest.raises(dea.DEAError):
        dea.host_resident_apply(s.row_ptr, s.col_ind, s.values, s.diag, s.b, s.b, 0.0, 3, 0)


def test_threaded_condensation_is_independent_of_the_thread_count(monkeypatch):
    """Above 2^17 rows the path factorisation runs on host threads (rows cut at anchors and
    path starts, merged in row order).  The reduced system is bit-identical with one thread
    and with many, and it condenses exactly: the expanded solution solves the full system."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    net = network.Network(network.config_params(4, 0, 8000))
    k = max(net.percolating_clusters(0), key=lambda q: net.cluster(q)[0])
    s = net.assemble(k, 1, 0, direction=0)
    assert s.n > 2 * 65536
    outs, cores = [], []
    for threads in ("1", "16"):
        monkeypatch.setenv("CICY_HOST_THREADS", threads)
        outs.append(dea.host_condense(s.row_ptr, s.col_ind, s.values, s.diag, s.b))
        cores.append(dea.host_core(s.row_ptr, s.col_ind, s.values, s.diag, s.b))
    monkeypatch.delenv("CICY_HOST_THREADS")
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    for a, b in zip(*cores):          # the series rounds too
        assert np.array_equal(a, b)
    anchors, rrp, rcol, rval, rdiag, rrhs = outs[1]
    m = len(anchors)
    rows = np.repeat(np.arange(m), np.diff(rrp))
    L = sp.csr_matrix((rval, (rows, rcol)), shape=(m, m))
    xj = sla.spsolve((L + L.T + sp.diags(rdiag)).tocsc(), rrhs)
    x = dea.host_expand(s.row_ptr, s.col_ind, s.values, s.diag, s.b, xj)
    ax = oracle.spmv_sss(s.row_ptr, s.col_ind, s.values, s.diag, x)
    assert np.linalg.norm(s.b - ax) <= 1e-11 * np.linalg.norm(s.b)


def floating_triangle_system(k=6, g=3.0):
    """The definite chain system plus a floating triangle of tube nodes (rows n0, n0 + 1,
    n0 + 2; conductance g on every side, no electrode): n0 and n0 + 2 are anchors (the side
    between them is a junction), n0 + 1 a path row.  The path elimination leaves the two
    anchors joined by 1.5 g, each with diagonal 1.5 g; the series round eliminates one, and
    the other is left without neighbours and with diagonal 1.5 g - (1.5 g)^2 / 1.5 g = 0."""
    rp0, col0, val0, diag0 = chain_system(k, g)
    n0 = len(diag0)
    rp, col, val = list(rp0), list(col0), list(val0)
    rows = {n0: [], n0 + 1: [(n0, -g)], n0 + 2: [(n0, -g), (n0 + 1, -g)]}
    for r in (n0, n0 + 1, n0 + 2):
        for c, v in rows[r]:
            col.append(c)
            val.append(v)
        rp.append(len(col))
    diag = np.concatenate([diag0, [2 * g] * 3])
    return ((rp0, col0, val0, diag0),
            (np.array(rp, np.int64), np.array(col, np.int64), np.array(val), diag), n0)


def test_floating_group_reduced_to_one_node_is_decoupled():
    """A floating group that the series rounds reduce to a single node leaves that node
    without neighbours and with a diagonal of rounding size (here exactly 0): the core gives
    it diagonal 1 instead of refusing the system, the group takes the voltage of its zero
    right-hand side (a solution of its singular, consistent block), and the definite part
    is solved exactly."""
    (rp0, col0, val0, diag0), (rp, col, val, diag), n0 = floating_triangle_system()
    b = np.zeros(len(diag))
    b[0] = 1.0
    rows, crp, ccol, cval, cdiag, crhs, rounds = dea.host_core(rp, col, val, diag, b)
    assert rounds >= 1
    q = int(np.searchsorted(rows, n0 + 2)) if n0 + 2 in set(rows) else int(
        np.searchsorted(rows, n0))
    assert rows[q] in (n0, n0 + 2)
    C = dense(crp, ccol, cval, cdiag)
    assert cdiag[q] == 1.0
    assert np.count_nonzero(C[q]) == 1 and np.count_nonzero(C[:, q]) == 1  # decoupled
    assert np.all(cdiag > 0.0)
    xc = np.linalg.solve(C, crhs)
    x = dea.host_core_expand(rp, col, val, diag, b, xc)
    A = dense(rp, col, val, diag)
    assert np.linalg.norm(b - A @ x) <= 1e-13
    assert np.allclose(x[n0:], x[n0], rtol=0, atol=1e-15)   # one voltage on the group
    want = np.linalg.solve(dense(rp0, col0, val0, diag0), b[:n0])
    assert np.allclose(x[:n0], want, rtol=1e-12)
