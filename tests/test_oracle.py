"""Pins of the CPU oracle (no GPU).

The reference ships no tests, golden vectors or fixtures for this path, so the oracle is
pinned by what the mathematics fixes:
  * closed forms: a uniform series chain (linear voltages, R = sum R_e), parallel chains
    (1/R = sum 1/R_k), a cubic lattice between two faces (linear in x,
    R = (nx-1)/(ny*nz) for unit resistors);
  * dense direct solves (numpy) on random conductance networks, committed as fixtures
    under tests/golden/ by tests/golden/make_fixtures.py;
  * invariants: current conservation between the two electrodes, Tellegen's identity
    V_app * I = sum_e g_e dV_e^2, the discrete maximum principle;
  * the SSS product against the dense symmetric matrix it stores.
A planted mistake (a dropped transpose term, a sign flip in the update) is shown to fail
the same checks.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from oracle import assembly

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dense(sys_):
    n = len(sys_["diag"])
    A = np.diag(sys_["diag"]).astype(float)
    rp, ci, v = sys_["row_ptr"], sys_["col_ind"], sys_["values"]
    for r in range(n):
        for j in range(rp[r], rp[r + 1]):
            A[r, ci[j]] += v[j]
            A[ci[j], r] += v[j]
    return A


def solve(sys_, tol=1e-10, x0=None, r0=None):
    n = len(sys_["diag"])
    x0 = np.zeros(n) if x0 is None else x0
    r0 = sys_["b"] if r0 is None else r0
    return oracle.pcg_sss(sys_["nodes"], sys_["reserved"], tol, sys_["row_ptr"],
                          sys_["col_ind"], sys_["values"], sys_["diag"], r0, sys_["vef"], x0)


def chain(gs, vapp=1.0):
    """Series chain: node 1 (V_app) - interior nodes - node 0 (0 V)."""
    m = len(gs) - 1          # interior nodes
    nodes = 2 + m
    edges = []
    seq = [1] + list(range(2, 2 + m)) + [0]
    for k, g in enumerate(gs):
        edges.append((seq[k], seq[k + 1], g))
    return assembly.build_from_edges(nodes, 2, edges, [0.0, vapp]), seq


def edge_current(volt, edges, node):
    return sum(g * (volt[j] - volt[i]) if i == node else g * (volt[i] - volt[j])
               for i, j, g in edges if node in (i, j))


def test_voltage_vector():
    assert list(oracle.voltage_vector(2, 10, 1, 2.5)) == [0.0, 2.5]
    assert list(oracle.voltage_vector(6, 10, 0, 1.0)) == [0, 10, 20, 30, 40, 50]
    with pytest.raises(ValueError):
        oracle.voltage_vector(2, 10, 2, 1.0)


def test_uniform_chain_is_linear():
    sys_, seq = chain([2.0] * 50)
    res = solve(sys_, tol=1e-13)
    assert res["status"] == oracle.ORC_OK
    volt = res["voltages"]
    # interior nodes 2..50 in series order from V=1 to V=0
    want = 1.0 - np.arange(1, 50) / 50.0
    assert np.allclose(volt[2:], want, atol=1e-11)
    # resistance = sum of series resistors
    I = oracle.boundary_current(np.array([0]), np.array([50]), np.array([0.5]), volt)
    assert abs(1.0 / I - 50 * 0.5) < 1e-9


def test_series_chain_voltages_follow_cumulative_resistance():
    rng = np.random.default_rng(0)
    gs = 10 ** rng.uniform(-6, -2, 30)
    sys_, seq = chain(list(gs))
    res = solve(sys_, tol=1e-14)
    R = 1 / gs
    cum = np.cumsum(R)[:-1]
    want = 1.0 - cum / R.sum()
    assert np.allclose(res["voltages"][2:], want, rtol=1e-8, atol=1e-10)


def test_parallel_chains():
    # three chains between node 1 (V=1) and node 0 (V=0)
    chains = [[1.0, 2.0, 4.0], [0.5] * 5, [3.0, 3.0]]
    edges, nxt = [], 2
    for gs in chains:
        prev = 1
        for g in gs[:-1]:
            edges.append((prev, nxt, g))
            prev, nxt = nxt, nxt + 1
        edges.append((prev, 0, gs[-1]))
    sys_ = assembly.build_from_edges(nxt, 2, edges, [0.0, 1.0])
    res = solve(sys_, tol=1e-14)
    I = edge_current(res["voltages"], edges, 0)
    Rk = [sum(1 / g for g in gs) for gs in chains]
    assert abs(1 / I - oracle.parallel_resistance(np.array(Rk))) < 1e-10
    assert abs(oracle.parallel_resistance(np.array(Rk)) - 1 / sum(1 / r for r in Rk)) < 1e-15


def lattice(nx, ny, nz, g=1.0):
    """Cubic lattice; face x=0 tied to node 1 (V=1), face x=nx-1 to node 0 (V=0)."""
    idx = {}
    nxt = 2
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if i == 0:
                    idx[i, j, k] = 1
                elif i == nx - 1:
                    idx[i, j, k] = 0
                else:
                    idx[i, j, k] = nxt
                    nxt += 1
    edges = []
    for (i, j, k), a in idx.items():
        for di, dj, dk in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            nb = (i + di, j + dj, k + dk)
            if nb in idx:
                b = idx[nb]
                if a < 2 and b < 2:
                    continue
                if a == b:
                    continue
                edges.append((a, b, g))
    return idx, edges, nxt


def test_cubic_lattice_closed_form():
    nx, ny, nz = 7, 4, 3
    idx, edges, nodes = lattice(nx, ny, nz)
    # boundary-boundary duplicates within a face are skipped above; lateral edges inside
    # the electrode faces connect equal nodes and carry no current
    dup = {}
    for a, b, g in edges:
        key = (max(a, b), min(a, b))
        dup[key] = dup.get(key, 0) + g
    merged = [(a, b, g) for (a, b), g in dup.items()]
    sys_ = assembly.build_from_edges(nodes, 2, merged, [0.0, 1.0])
    res = solve(sys_, tol=1e-14)
    v = res["voltages"]
    for (i, j, k), a in idx.items():
        assert abs(v[a] - (1.0 - i / (nx - 1))) < 1e-10
    I = edge_current(v, merged, 0)
    assert abs(1 / I - (nx - 1) / (ny * nz)) < 1e-10


def test_spmv_matches_dense_and_catches_dropped_transpose():
    rng = np.random.default_rng(5)
    n = 60
    edges = [(i, i - 1, rng.uniform(0.1, 1)) for i in range(3, n)]
    edges += [(int(a), int(b), rng.uniform(0.1, 1)) for a, b in rng.integers(2, n, (80, 2)) if a != b]
    edges += [(2, 1, 0.7), (n - 1, 0, 0.4)]
    sys_ = assembly.build_from_edges(n, 2, edges, [0.0, 1.0])
    v = rng.standard_normal(n - 2)
    A = dense(sys_)
    assert np.allclose(oracle.spmv_sss(sys_["row_ptr"], sys_["col_ind"], sys_["values"],
                                       sys_["diag"], v), A @ v, rtol=1e-13, atol=1e-13)
    # planted mistake: the transpose half dropped (row gathers only)
    wrong = sys_["diag"] * v
    for r in range(n - 2):
        for j in range(sys_["row_ptr"][r], sys_["row_ptr"][r + 1]):
            wrong[r] += sys_["values"][j] * v[sys_["col_ind"][j]]
    assert not np.allclose(wrong, A @ v, rtol=1e-6)


def test_update_residual():
    rng = np.random.default_rng(7)
    sys_, _ = chain(list(rng.uniform(0.5, 2, 20)))
    x0 = rng.uniform(0, 1, len(sys_["diag"]))
    r = oracle.update_residual(sys_["row_ptr"], sys_["col_ind"], sys_["values"], sys_["diag"],
                               x0, sys_["b"])
    assert np.allclose(r, sys_["b"] - dense(sys_) @ x0, atol=1e-14)


def test_pcg_stops_relative_to_initial_residual_and_counts_iterations():
    rng = np.random.default_rng(9)
    sys_, _ = chain(list(10 ** rng.uniform(-3, 0, 40)))
    n = len(sys_["diag"])
    x0 = np.linspace(1, 0, n)
    r0 = oracle.update_residual(sys_["row_ptr"], sys_["col_ind"], sys_["values"],
                                sys_["diag"], x0, sys_["b"])
    res = solve(sys_, tol=1e-10, x0=x0, r0=r0)
    assert res["status"] == oracle.ORC_OK
    assert np.linalg.norm(res["r"]) <= 1e-10 * np.linalg.norm(r0)
    # one iteration less does not meet the rule (k is the first iteration that does)
    res2 = oracle.pcg_sss(sys_["nodes"], 2, 1e-10, sys_["row_ptr"], sys_["col_ind"],
                          sys_["values"], sys_["diag"], r0, sys_["vef"], x0,
                          max_iter=res["iters"] - 1)
    assert res2["status"] == oracle.ORC_MAXITER
    assert np.linalg.norm(res2["r"]) > 1e-10 * np.linalg.norm(r0)
    # the recursive residual stays close to the true one
    tr = sys_["b"] - dense(sys_) @ res["x"]
    assert np.linalg.norm(tr - res["r"]) <= 1e-8 * np.linalg.norm(sys_["b"])


def test_pcg_planted_sign_error_fails():
    """The same chain solved with the residual update sign flipped does not converge to
    the closed form -- the closed-form check catches that kind of slip."""
    sys_, _ = chain([1.0] * 20)
    A = dense(sys_)
    b = sys_["b"]
    x = np.zeros(len(b)); r = b.copy(); z = r / np.diag(A); p = z.copy(); rz = r @ z
    for _ in range(40):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r += a * q          # planted: should be r -= a q
        z = r / np.diag(A)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    want = 1.0 - np.arange(1, 20) / 20.0
    assert not np.allclose(x, want, atol=1e-6)


def test_tellegen_and_maximum_principle_on_random_network():
    rng = np.random.default_rng(11)
    n = 200
    edges = [(i, i - 1, 10 ** rng.uniform(-4, 0)) for i in range(3, n)]
    edges += [(int(a), int(b), 10 ** rng.uniform(-7, 0)) for a, b in rng.integers(2, n, (300, 2))
              if a != b]
    edges += [(2, 1, 1.0), (n - 1, 0, 1.0), (n // 2, 1, 1e-3)]
    merged = {}
    for a, b, g in edges:
        merged[(max(a, b), min(a, b))] = g    # overwrite rule
    edges = [(a, b, g) for (a, b), g in merged.items()]
    sys_ = assembly.build_from_edges(n, 2, edges, [0.0, 1.0])
    res = solve(sys_, tol=1e-13)
    v = res["voltages"]
    assert v[2:].min() >= -1e-9 and v[2:].max() <= 1 + 1e-9
    I0 = edge_current(v, edges, 0)     # into node 0
    I1 = edge_current(v, edges, 1)     # out of node 1
    assert abs(I0 + I1) <= 1e-8 * abs(I0)
    power = sum(g * (v[a] - v[b]) ** 2 for a, b, g in edges)
    assert abs(power - 1.0 * abs(I0)) <= 1e-8 * power


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "dense_*.npz"))))
def test_golden_dense_fixtures(path):
    f = np.load(path, allow_pickle=False)
    sys_ = {k: f[k] for k in ("row_ptr", "col_ind", "values", "diag", "b", "vef")}
    sys_["nodes"] = int(f["nodes"])
    sys_["reserved"] = int(f["reserved"])
    res = solve(sys_, tol=1e-12)
    x_true = f["x"]
    err = np.linalg.norm(res["x"] - x_true) / np.linalg.norm(x_true)
    assert res["status"] == oracle.ORC_OK and err < 1e-8
    xt, _, rel = oracle.cg_truth(sys_["row_ptr"], sys_["col_ind"], sys_["values"],
                                 sys_["diag"], sys_["b"], tol=1e-14)
    assert np.linalg.norm(xt - x_true) / np.linalg.norm(x_true) < 1e-10


def test_goldens_exist():
    assert len(glob.glob(os.path.join(GOLDEN, "dense_*.npz"))) >= 3


def test_element_laws():
    c1, c2 = oracle.tunnel_constants()
    s = np.sqrt(2.0 * 9.10938291 * 2.0 * 1.60217662)
    assert abs(c1 - 10.0 * 6.62606957 ** 2 / (9.10938291 ** 2 * s)) < 1e-15
    assert abs(c2 - 4000.0 * np.pi * s / 6.62606957) < 1e-9
    assert abs(c1 - 0.6924754759963906) < 1e-12 and abs(c2 - 14490.505704298546) < 1e-8
    # straight 20 nm segment of a 10 nm tube at 20 Ohm.um
    xyz = np.array([[0.0, 0, 0], [0.02, 0, 0]])
    assert abs(oracle.resistance_cnt(xyz, 0, 1, 0.005, 20.0) - 20 * 0.02 / (np.pi * 0.005 ** 2)) < 1e-9
    # polyline length adds up
    xyz = np.array([[0.0, 0, 0], [0.03, 0.04, 0], [0.03, 0.04, 0.12]])
    assert abs(oracle.resistance_cnt(xyz, 0, 2, 0.005, 20.0) -
               20 * 0.17 / (np.pi * 0.005 ** 2)) < 1e-9
    # junction: area from the smaller diameter
    d = 0.00034
    want = c1 * d * np.exp(c2 * d) / (0.01 ** 2)
    assert abs(oracle.junction_resistance(d, 0.005, 1, 0.005, 1, c1, c2) - want) < 1e-9 * want
    assert abs(oracle.junction_resistance(d, 0.005, 1, 0.01, 1, c1, c2) - want) < 1e-9 * want
