"""GPU: the device assembly (cicy_dnet_assemble) against the host assembler
(cicy_net_assemble, which the CPU tests check bit for bit against the dictionary restatement
of the reference's std::map assembly, oracle/assembly.py), against hand-computed systems,
and end to end into the solver.

Node numbers, row_ptr, col_ind, the boundary lists and the plate groups must be identical.
The values, diagonal, b, x0 and r0 are identical with unit resistors and with the constant
junction law; with the tunnelling law they may differ where a junction's exp() enters: both
libraries' exp are within an ulp of the exact value, so the bar there is a few ulps."""
import ctypes

import numpy as np
import pytest

from cicy_geom_amd import dea, network

pytestmark = pytest.mark.gpu

ULP_BAR = 8 * np.finfo(np.float64).eps  # relative, per entry, tunnelling law only


@pytest.fixture(scope="module")
def ctx():
    c = dea.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def nets():
    out = {"c1": network.generate(1),
           "plates": network.Network(network.NetworkParams(box=2.0, n_cnt=2560, n_gnp=31,
                                                           seed=5))}
    yield out
    for n in out.values():
        n.close()


def largest(net, direction=0):
    ks = net.percolating_clusters(direction)
    return max(ks, key=lambda c: net.cluster(c)[0])


def rel_diff(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), np.finfo(np.float64).tiny)))


def compare(dev, s, exact, blocks=None):
    """Device assembly `dev` against host assembly `s`."""
    assert (dev.nodes, dev.reserved, dev.n, dev.nnz) == (s.nodes, s.reserved, s.n, s.nnz)
    assert dev.has_guess == s.has_guess
    assert np.array_equal(dev.vef, s.vef)
    assert np.array_equal(dev.get("row_ptr"), s.row_ptr)
    assert np.array_equal(dev.get("col_ind"), s.col_ind)
    assert np.array_equal(dev.get("lmm"), s.lmm)
    if s.lmm_gnp is not None and len(s.lmm_gnp):
        assert np.array_equal(dev.get("lmm_gnp"), s.lmm_gnp)
    assert dev.gnp_nodes == s.gnp_nodes
    worst = 0.0
    for name in ("values", "diag", "b", "x0", "r0"):
        got, want = dev.get(name), getattr(s, name)
        if exact:
            assert np.array_equal(got, want), name
        elif name == "r0":
            # b - A x0 cancels: the bar is the size of the terms it sums
            rows = np.repeat(np.arange(s.n), np.diff(s.row_ptr))
            terms = np.abs(s.b) + np.abs(s.diag * s.x0)
            np.add.at(terms, rows, np.abs(s.values * s.x0[s.col_ind]))
            np.add.at(terms, s.col_ind, np.abs(s.values * s.x0[rows]))
            assert np.all(np.abs(got - want) <= 2 * ULP_BAR * terms), name
        else:
            d = rel_diff(got, want)
            worst = max(worst, d)
            assert d <= ULP_BAR, (name, d)
    if s.direction >= 0:
        assert dev.direction == s.direction
        for (eb, ei, er), (hb, hi, hr) in zip(dev.boundary_elements(), s.boundary_elements):
            assert np.array_equal(eb, hb) and np.array_equal(ei, hi)
            assert np.array_equal(er, hr)
    if blocks is not None:
        assert np.array_equal(dev.blocks, blocks)
    return worst


@pytest.mark.parametrize("name", ["c1", "plates"])
@pytest.mark.parametrize("r_flag,family,direction", [(1, 0, 0), (0, 0, 0), (1, 1, 1),
                                                     (1, 2, 2), (1, 3, 0), (0, 6, 1),
                                                     (1, 6, 2)])
def test_device_assembly_equals_host(ctx, nets, name, r_flag, family, direction):
    net = nets[name]
    k = largest(net)
    dn = net.device(ctx)
    try:
        for exp_law in (True, False):
            elec = network.ElectricParams(junction_exponential=exp_law)
            s = net.assemble(k, r_flag, family, elec, direction=direction)
            dev = dn.assemble(k, r_flag, family, elec, direction)
            blocks = network.plate_row_ranges(net, k, s) if family == 0 else None
            worst = compare(dev, s, exact=(r_flag == 0 or not exp_law), blocks=blocks)
            print(f"\n{name} r_flag {r_flag} family {family} exp {exp_law}: n {s.n} nnz {s.nnz}"
                  f" worst rel diff {worst:.2e}")
            dev.close()
    finally:
        dn.close()


def test_every_cluster_of_a_network(ctx, nets):
    """Every cluster (percolating or not) of the plate network assembles as on the host,
    with the constant junction law (bit for bit)."""
    net = nets["plates"]
    elec = network.ElectricParams(junction_exponential=False)
    dn = net.device(ctx)
    try:
        for k in range(net.count("clusters")):
            s = net.assemble(k, 1, 0, elec, direction=0)
            dev = dn.assemble(k, 1, 0, elec, 0)
            compare(dev, s, exact=True)
            dev.close()
    finally:
        dn.close()


def solve_x(sysd, blocks, method):
    if method == dea.METHOD_CONDENSED_BLOCK:
        sysd.set_blocks(blocks)
    st = sysd.solve(method, 1e-10, 0, dea.REL_B)
    x, _ = sysd.get()
    return st, x


@pytest.mark.parametrize("method", [dea.METHOD_CONDENSED_BLOCK, dea.METHOD_PCG_F64])
def test_device_assembly_into_the_solver(ctx, nets, method):
    """Assembled on the device and solved without a copy through the host: with the constant
    junction law the solve is the host-assembled system's solve bit for bit (same iterations,
    same x, same boundary currents); with the tunnelling law the device solution solves the
    host-assembled system to the same residual bar."""
    net = nets["plates"]
    k = largest(net)
    dn = net.device(ctx)
    try:
        for exp_law in (False, True):
            elec = network.ElectricParams(junction_exponential=exp_law)
            s = net.assemble(k, 1, 0, elec, direction=0)
            blocks = network.plate_row_ranges(net, k, s)
            dev = dn.assemble(k, 1, 0, elec, 0)
            sd = dev.system()
            st_d, x_d = solve_x(sd, dev.blocks, method)
            sd.close()
            sh = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
            sh.set_rhs(s.r0, s.x0 if s.has_guess else None, s.b)
            st_h, x_h = solve_x(sh, blocks, method)
            sh.close()
            assert st_d.status == dea.CICY_OK and st_h.status == dea.CICY_OK
            volt_d = np.concatenate([dev.vef, x_d])
            volt_h = np.concatenate([s.vef, x_h])
            cur_d = [ctx.boundary_current(*e, volt_d) for e in dev.boundary_elements()]
            cur_h = [ctx.boundary_current(*e, volt_h) for e in s.boundary_elements]
            if not exp_law:
                assert st_d.iters == st_h.iters
                assert np.array_equal(x_d, x_h)
                assert cur_d == cur_h
            else:
                # the device system's values are within a few ulps of the host's, so the
                # device solution solves the host-assembled system to the same residual bar
                # (plus the rounding of the product)
                rows = np.repeat(np.arange(s.n), np.diff(s.row_ptr))
                ax = s.diag * x_d
                np.add.at(ax, rows, s.values * x_d[s.col_ind])
                np.add.at(ax, s.col_ind, s.values * x_d[rows])
                mag = np.abs(s.diag * x_d)
                np.add.at(mag, rows, np.abs(s.values * x_d[s.col_ind]))
                np.add.at(mag, s.col_ind, np.abs(s.values * x_d[rows]))
                res = np.linalg.norm(s.b - ax)
                assert res <= 1e-10 * np.linalg.norm(s.b) * 1.01 + 16 * ULP_BAR * np.linalg.norm(mag)
                assert np.isfinite(cur_d).all()
            dev.close()
    finally:
        dn.close()


# ---- hand-built networks ------------------------------------------------------------------

class Flat:
    """A cicy_net_arrays over numpy arrays (kept alive here)."""

    def __init__(self, xyz, radius, elements, junctions=(), faces_pts=None, faces_tubes=None,
                 clusters=None, box=1.0):
        self.keep = []
        a = dea.NetArrays()
        xyz = np.ascontiguousarray(xyz, np.float64)
        a.n_points = len(xyz)
        a.xyz = self._p(xyz)
        a.n_tubes = len(radius)
        a.radius = self._p(np.asarray(radius, np.float64))
        ep, pts = self._lists(elements)
        a.elem_ptr, a.elem_pts = self._p(ep), self._p(pts)
        jr = np.zeros(len(junctions), dtype=[("p1", "i8"), ("p2", "i8"), ("e1", "i8"),
                                             ("e2", "i8"), ("dist", "f8")])
        for i, j in enumerate(junctions):
            jr[i] = j
        a.n_junctions = len(junctions)
        a.junctions = self._p(jr)
        a.n_gnp_points = a.n_gnps = a.n_mixed = a.n_gjunctions = 0
        empty_g = np.zeros(1, np.int64)
        a.gstruct_ptr = a.gtri_ptr = self._p(empty_g)
        for nm, lists in (("bpts", faces_pts or [[]] * 6), ("bcnt", faces_tubes or [[]] * 6),
                          ("gbpts", [[]] * 6), ("gbnd", [[]] * 6)):
            p, e = self._lists(lists)
            setattr(a, nm + "_ptr", self._p(p))
            setattr(a, nm, self._p(e))
        clusters = clusters or [dict(tubes=list(range(len(radius))),
                                     junc=list(range(len(junctions))))]
        a.n_clusters = len(clusters)
        for nm, key in (("cl_tubes", "tubes"), ("cl_junc", "junc"), ("cl_gnps", "gnps"),
                        ("cl_mixed", "mixed"), ("cl_gjunc", "gjunc")):
            p, e = self._lists([c.get(key, []) for c in clusters])
            setattr(a, nm + "_ptr", self._p(p))
            setattr(a, nm, self._p(e))
        for d in range(3):
            a.window_lo[d] = 0.0
            a.window_hi[d] = box
        self.a = a

    def _p(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.size == 0:
            arr = np.zeros(1, arr.dtype)
        self.keep.append(arr)
        return arr.ctypes.data

    def _lists(self, ll):
        ptr = np.zeros(len(ll) + 1, np.int64)
        ptr[1:] = np.cumsum([len(x) for x in ll])
        ent = np.concatenate([np.asarray(x, np.int64) for x in ll]) if ll else np.zeros(0)
        return ptr, np.asarray(ent, np.int64)


def faces(**kw):
    out = [[] for _ in range(6)]
    for k, v in kw.items():
        out[int(k[1:])] = v
    return out


RHO, RAD, V = 20.0, 0.005, 1.0


def seg_g(length):
    return 1.0 / (RHO * length / (np.pi * RAD * RAD))


def test_hand_built_single_tube(ctx):
    """One straight tube from face 2 to face 4 through x = 0.25 and 0.5 (element points 0,
    2, 3 of four polyline points): reserved nodes 0 and 1 for the faces, free nodes 2 and 3;
    Calculate_resistance_cnt sums the polyline between element points; b, the ramp guess
    V/L (max - x) and r0 = b - A x0 follow by hand, and both ends are boundary elements of
    either face's list (the host's rule: an end whose node is 0 or 1)."""
    xyz = [[0.0, 0.5, 0.5], [0.125, 0.5, 0.5], [0.25, 0.5, 0.5], [1.0, 0.5, 0.5]]
    f = Flat(xyz, [RAD], [[0, 2, 3]], faces_pts=faces(f2=[0], f4=[3]),
             faces_tubes=faces(f2=[0], f4=[0]))
    dn = ctx.device_network(f.a)
    dev = dn.assemble(0, 1, 0, network.ElectricParams(applied_voltage=V, resistivity_cnt=RHO),
                      0)
    assert (dev.nodes, dev.reserved, dev.n, dev.nnz) == (3, 2, 1, 0)
    assert np.array_equal(dev.get("lmm"), [0, -1, 2, 1])
    g1 = 1.0 / (RHO * (0.125 + 0.125) / (np.pi * RAD * RAD))
    g2 = seg_g(0.75)
    assert dev.get("diag")[0] == pytest.approx(g1 + g2, rel=1e-15)
    assert dev.get("b")[0] == pytest.approx(g2 * V, rel=1e-15)
    assert dev.has_guess
    x0 = V / 1.0 * (1.0 - 0.25)
    assert dev.get("x0")[0] == pytest.approx(x0, rel=1e-15)
    assert dev.get("r0")[0] == pytest.approx(g2 * V - (g1 + g2) * x0, rel=1e-12)
    for w, (eb, ei, er) in enumerate(dev.boundary_elements()):
        assert np.array_equal(eb, [0, 1]) and np.array_equal(ei, [2, 2])
        assert er == pytest.approx([1 / g1, 1 / g2], rel=1e-15)
    dev.close()
    dn.close()


def test_hand_built_three_point_element_and_junction(ctx):
    """Tube 0 has three element points whose ends both sit on face 2 (reserved node 0): the
    reference's Three_points_in_element gives one resistor g1 + g2 between node 0 and the
    middle node.  A junction joins that middle point to tube 1, which runs to face 4: the
    junction enters the middle node's diagonal and tube 1's first row."""
    xyz = [[0.0, 0.2, 0.5], [0.3, 0.2, 0.5], [0.0, 0.4, 0.5],      # tube 0 (a hairpin)
           [0.3, 0.2, 0.502], [0.6, 0.2, 0.5], [1.0, 0.2, 0.5]]   # tube 1
    jd = 0.001
    f = Flat(xyz, [RAD, RAD], [[0, 1, 2], [3, 4, 5]], junctions=[(1, 3, 0, 1, jd)],
             faces_pts=faces(f2=[0, 2], f4=[5]), faces_tubes=faces(f2=[0], f4=[1]))
    elec = network.ElectricParams(applied_voltage=V, resistivity_cnt=RHO,
                                  junction_exponential=False, junction_resistance=1e4)
    dn = ctx.device_network(f.a)
    dev = dn.assemble(0, 1, 0, elec, 0)
    # nodes: reserved 0 (face 2), 1 (face 4); tube 0 middle point 2; tube 1: points 3, 4
    assert np.array_equal(dev.get("lmm"), [0, 2, 0, 3, 4, 1])
    assert (dev.n, dev.nnz) == (3, 2)
    ga = seg_g(0.3)
    gb = seg_g(np.hypot(0.3, 0.2))
    gj = 1.0 / 1e4
    g34 = seg_g(np.hypot(0.3, 0.002))
    g41 = seg_g(0.4)
    assert np.array_equal(dev.get("row_ptr"), [0, 0, 1, 2])
    assert np.array_equal(dev.get("col_ind"), [0, 1])
    vals = dev.get("values")
    assert vals[0] == pytest.approx(-gj, rel=1e-15)
    assert vals[1] == pytest.approx(-g34, rel=1e-15)
    assert dev.get("diag") == pytest.approx([ga + gb + gj, gj + g34, g34 + g41], rel=1e-14)
    assert dev.get("b") == pytest.approx([0.0, 0.0, g41 * V], rel=1e-15)
    dev.close()
    dn.close()


def test_hand_built_errors(ctx):
    """The reference's error conditions give its messages: a tube with one element point
    (LM_matrix_for_cnts), a free node whose diagonal is zero (Fill_sparse_stiffness_matrix;
    here every resistor has zero conductance because the tube resistivity is infinite).
    Malformed networks are refused when copied: out-of-range indices, element points out of
    order, a point that is an element point of two tubes."""
    xyz = [[0.0, 0.5, 0.5], [0.5, 0.5, 0.5], [1.0, 0.5, 0.5], [0.7, 0.7, 0.7]]
    f = Flat(xyz, [RAD, RAD], [[0, 1, 2], [3]], faces_pts=faces(f2=[0], f4=[2]))
    dn = ctx.device_network(f.a)
    with pytest.raises(dea.DEAError, match="a tube with not enough points"):
        dn.assemble(0, 1, 0, None, 0)
    dn.close()
    f = Flat(xyz[:3], [RAD], [[0, 1, 2]], faces_pts=faces(f2=[0], f4=[2]))
    dn = ctx.device_network(f.a)
    dev = dn.assemble(0, 1, 0, None, 0)  # the same tube alone assembles
    assert (dev.n, dev.nnz) == (1, 0)
    dev.close()
    with pytest.raises(dea.DEAError, match="zero element on the diagonal"):
        dn.assemble(0, 1, 0, network.ElectricParams(resistivity_cnt=np.inf), 0)
    with pytest.raises(dea.DEAError, match="invalid family"):
        dn.assemble(0, 1, 7, None, 0)
    with pytest.raises(dea.DEAError, match="cluster index"):
        dn.assemble(1, 1, 0, None, 0)
    dn.close()
    bad = Flat(xyz, [RAD], [[0, 1, 7]])
    with pytest.raises(dea.DEAError, match=r"elem_pts\[2\] = 7 is outside"):
        ctx.device_network(bad.a)
    bad = Flat(xyz, [RAD], [[0, 2, 1]])
    with pytest.raises(dea.DEAError, match="not ascending"):
        ctx.device_network(bad.a)
    bad = Flat(xyz, [RAD, RAD], [[0, 1, 2], [1, 3]])
    with pytest.raises(dea.DEAError, match="element point of two tubes"):
        ctx.device_network(bad.a)
    bad = Flat(xyz, [RAD], [[0, 1, 2]], junctions=[(0, 2, 0, 5, 0.001)])
    with pytest.raises(dea.DEAError, match=r"junctions\[0\] has an index out of range"):
        ctx.device_network(bad.a)


def test_config3_device_assembly(ctx):
    """The benchmark's network (8 um RVE, 160k tubes + 2k plates): the largest x-percolating
    cluster's ~6M-row system assembles on the device as on the host (structure and boundary
    lists identical, values within the tunnelling-law bar), and the plate groups are the
    ones the block method uses."""
    net = network.generate(3)
    k = largest(net)
    s = net.assemble(k, 1, 0, direction=0)
    blocks = network.plate_row_ranges(net, k, s)
    dn = net.device(ctx)
    dev = dn.assemble(k, 1, 0, None, 0)
    worst = compare(dev, s, exact=False, blocks=blocks)
    print(f"\nconfig 3: n {s.n} nnz {s.nnz} blocks {len(blocks)} worst rel diff {worst:.2e}")
    dev.close()
    dn.close()
    net.close()
