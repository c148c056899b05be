"""Host form of the block-preconditioned core CG (CICY_METHOD_CONDENSED_BLOCK): the layout
(plate blocks, dense inverses) and the Chronopoulos-Gear recurrences the device kernel
runs, checked against a sparse direct solve.  No GPU."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from cicy_geom_amd import dea, network


def full_matrix(row_ptr, col_ind, values, diag):
    n = len(diag)
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    L = sp.csr_matrix((values, (rows, col_ind)), shape=(n, n))
    return (L + L.T + sp.diags(diag)).tocsc()


@pytest.fixture(scope="module")
def plate_system():
    # 2 um RVE at the tube density of the 5 um / 40k-tube workload and the plate density of
    # the 8 um / 2k-plate one
    p = network.NetworkParams(box=2.0, n_cnt=2560, n_gnp=31, seed=5)
    net = network.Network(p)
    k = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    s = net.assemble(k, 1, 0, direction=0)
    ranges = network.plate_row_ranges(net, k, s)
    net.close()
    return s, ranges


def core_ranges(rows, ranges):
    out = []
    for a, b in ranges:
        q0, q1 = np.searchsorted(rows, a), np.searchsorted(rows, b)
        if q1 - q0 >= 2:
            out.append((q0, q1))
    return np.array(out, np.int64).reshape(-1, 2)


def test_plate_ranges_cover_the_plate_nodes(plate_system):
    s, ranges = plate_system
    assert len(ranges) >= 10
    covered = int(np.sum(ranges[:, 1] - ranges[:, 0]))
    assert covered == s.gnp_nodes
    assert np.all(ranges[1:, 0] >= ranges[:-1, 1])
    assert ranges[-1, 1] == s.n  # plate nodes are numbered after every tube node


def test_plate_blocks_cut_iterations_and_keep_the_solution(plate_system):
    s, ranges = plate_system
    rows, crp, ccol, cval, cdiag, crhs, _ = dea.host_core(s.row_ptr, s.col_ind, s.values,
                                                         s.diag, s.b)
    ref2 = float(s.b @ s.b)
    A = full_matrix(s.row_ptr, s.col_ind, s.values, s.diag)
    xt = sla.spsolve(A, s.b)
    out = {}
    for name, rr in (("jacobi", np.zeros((0, 2), np.int64)), ("plates", core_ranges(rows, ranges))):
        xc, it, br = dea.host_core_block_cg(crp, ccol, cval, cdiag, crhs, rr, 1e-10, ref2)
        x = dea.host_core_expand(s.row_ptr, s.col_ind, s.values, s.diag, s.b, xc)
        res = np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b)
        err = np.linalg.norm(x - xt) / np.linalg.norm(xt)
        out[name] = (it, br, res, err)
        print(name, out[name])
        assert res <= 1e-10 * (1 + 1e-6)
    # the same stopping rule, ten times fewer iterations with exact plate blocks
    assert out["plates"][1] > 0 and out["jacobi"][1] == 0
    assert out["plates"][0] * 10 < out["jacobi"][0]
    # no worse than the reference algorithm's own solve at the same residual
    ref = oracle.pcg_sss(s.nodes, s.reserved, 1e-10 * np.linalg.norm(s.b) / np.linalg.norm(s.r0),
                         s.row_ptr, s.col_ind, s.values, s.diag, s.r0, s.vef, s.x0)
    err_ref = np.linalg.norm(ref["x"] - xt) / np.linalg.norm(xt)
    assert out["plates"][3] <= max(3 * err_ref, 1e-9), (out["plates"][3], err_ref)


def test_block_inverse_is_exact_on_a_single_block():
    """One dense block covering a whole SPD system: the preconditioned operator is the
    identity up to the FP32 rounding of the inverse's remainder, so CG converges in one
    iteration to that rounding and in two to 1e-12."""
    rng = np.random.default_rng(3)
    n = 40
    G = np.abs(rng.normal(size=(n, n)))
    G = np.triu(G, 1)
    G = G + G.T
    diag = G.sum(axis=1) + rng.uniform(0.1, 1.0, n)
    row_ptr, col_ind, values = [0], [], []
    for i in range(n):
        for j in range(i):
            col_ind.append(j)
            values.append(-G[i, j])
        row_ptr.append(len(col_ind))
    b = rng.normal(size=n)
    x, it, br = dea.host_core_block_cg(np.array(row_ptr), np.array(col_ind), np.array(values),
                                       diag, b, [(0, n)], 1e-12)
    A = np.diag(diag) - G
    assert br == n and it <= 2
    _, it1, _ = dea.host_core_block_cg(np.array(row_ptr), np.array(col_ind), np.array(values),
                                       diag, b, [(0, n)], 1e-6)
    assert it1 == 1
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_overlapping_groups_are_rejected():
    n = 5
    row_ptr = np.array([0, 0, 1, 2, 3, 4])
    col_ind = np.array([0, 1, 2, 3])
    values = -np.ones(4)
    diag = np.full(n, 3.0)
    with pytest.raises(dea.DEAError):
        dea.host_core_block_cg(row_ptr, col_ind, values, diag, np.ones(n), [(0, 3), (2, 5)])
