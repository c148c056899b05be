"""Writes the golden fixtures of tests/test_oracle.py and tests/test_assembly.py.

Run from the repository root:  python tests/golden/make_fixtures.py

dense_*.npz : random conductance networks (chains plus random junctions, conductances
              spread over several decades, 2/4/6 Dirichlet nodes) assembled with the
              reference's rules and solved by a dense LU solve (numpy.linalg.solve).
c1_solution.npz : the x-direction resistance system of the 1 um / 300-tube workload
              (seed 1_000_000), solved by a sparse direct LU solve (scipy splu), with
              a SHA-256 of the assembled arrays so generator drift is detected.
No reference code is run: the reference ships no tests or vectors for this path.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import assembly  # noqa: E402


def dense_matrix(s):
    n = len(s["diag"])
    A = np.diag(s["diag"])
    for r in range(n):
        for j in range(s["row_ptr"][r], s["row_ptr"][r + 1]):
            A[r, s["col_ind"][j]] += s["values"][j]
            A[s["col_ind"][j], r] += s["values"][j]
    return A


def random_network(seed, n, reserved, spread):
    rng = np.random.default_rng(seed)
    edges = {}
    for i in range(reserved + 1, n):
        edges[(i, i - 1)] = 10 ** rng.uniform(-spread / 2, 0)
    for a, b in rng.integers(reserved, n, (n, 2)):
        if a != b:
            edges[(max(a, b), min(a, b))] = 10 ** rng.uniform(-spread, 0)
    for d in range(reserved):
        for t in rng.integers(reserved, n, 3):
            edges[(int(t), d)] = 10 ** rng.uniform(-2, 0)
    vef = [0.0]
    for _ in range(1, reserved):
        vef.append(vef[-1] + 1.0)
    return assembly.build_from_edges(n, reserved, [(a, b, g) for (a, b), g in edges.items()], vef)


def system_hash(s):
    h = hashlib.sha256()
    for k in ("row_ptr", "col_ind", "values", "diag", "b"):
        h.update(np.ascontiguousarray(getattr(s, k) if not isinstance(s, dict) else s[k]).tobytes())
    return h.hexdigest()


def main():
    cases = [(1, 60, 2, 3.0), (2, 300, 2, 7.0), (3, 400, 4, 5.0), (4, 500, 6, 7.0),
             (5, 800, 2, 9.0)]
    for seed, n, rd, spread in cases:
        s = random_network(seed, n, rd, spread)
        x = np.linalg.solve(dense_matrix(s), s["b"])
        np.savez_compressed(os.path.join(HERE, f"dense_{seed}.npz"), nodes=s["nodes"],
                            reserved=s["reserved"], row_ptr=s["row_ptr"], col_ind=s["col_ind"],
                            values=s["values"], diag=s["diag"], b=s["b"], vef=s["vef"], x=x)
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from cicy_geom_amd import network
    net = network.generate(1)
    s = network.largest_percolating_system(net, 0)
    n = s.n
    rows = np.repeat(np.arange(n), np.diff(s.row_ptr))
    L = sp.coo_matrix((s.values, (rows, s.col_ind)), shape=(n, n))
    A = (L + L.T + sp.diags(s.diag)).tocsc()
    x = sla.splu(A).solve(s.b)
    np.savez_compressed(os.path.join(HERE, "c1_solution.npz"), x=x, sha256=system_hash(s),
                        n=n, nnz=s.nnz)
    print("fixtures written")


if __name__ == "__main__":
    main()
