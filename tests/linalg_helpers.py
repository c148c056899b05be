"""Shared checks for the GPU parity tests: a refined direct solve (the truth), the rows of
components tied to an electrode, and the current's resolution.  Test infrastructure."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csgraph
import scipy.sparse.linalg as sla


def full_matrix(s):
    """A = L + L^T + D of an assembled system (scipy CSC)."""
    rows = np.repeat(np.arange(s.n), np.diff(s.row_ptr))
    L = sp.csr_matrix((s.values, (rows, s.col_ind)), shape=(s.n, s.n))
    return (L + L.T + sp.diags(s.diag)).tocsc()


def grounded_rows(s):
    """Rows in a connected component of the free-node graph that couples to an electrode
    (a row of A with a positive sum); the others float: a tube group whose only junction to
    the rest reached an electrode and was dropped (Direct_Electrifying.cpp:817-837) has a
    singular, consistent block on which CG keeps the mean of its initial guess."""
    rows = np.repeat(np.arange(s.n), np.diff(s.row_ptr))
    G = sp.csr_matrix((np.ones(len(rows)), (rows, s.col_ind)), shape=(s.n, s.n))
    ncomp, lab = csgraph.connected_components(G + G.T, directed=False)
    rs = s.diag.copy()
    np.add.at(rs, rows, s.values)
    np.add.at(rs, s.col_ind, s.values)
    tied = np.zeros(ncomp, bool)
    tied[lab[rs > 1e-9 * s.diag]] = True
    return tied[lab]


def refined_truth(s):
    """LU solve of the free block, refined three times with b - A x summed in extended
    precision (until it stops moving)."""
    A = full_matrix(s)
    lu = sla.splu(A)
    x = lu.solve(s.b)
    Ac = A.tocoo()
    data = Ac.data.astype(np.longdouble)
    for _ in range(3):
        ax = np.zeros(s.n, np.longdouble)
        np.add.at(ax, Ac.row, data * x.astype(np.longdouble)[Ac.col])
        x = x + lu.solve((s.b.astype(np.longdouble) - ax).astype(np.float64))
    return x


def smallest_eigenvalue(s, rows):
    """The smallest eigenvalue of A restricted to `rows` (shift-invert Lanczos)."""
    A = full_matrix(s)[rows][:, rows].tocsc()
    return float(sla.eigsh(A, k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0])


def resolution_floor(s, v, I):
    """eps * max|V| * sum_e g_e / |I|: the relative change of the boundary current
    I = sum_e g_e (V_int - V_bnd) when every voltage moves by one unit in the last place of
    the largest voltage.  A converged FP64 solve cannot determine the voltages better than a
    few such units, so this is how closely FP64 voltages determine R.  Where a plate's
    triangulation edges (up to ~1e4 S) touch an electrode it is 1e-10 and above; on tube
    networks it is near 1e-14."""
    g = 1.0 / s.boundary_elements[0][2]
    return np.finfo(float).eps * np.abs(v).max() * g.sum() / abs(I)
