"""Nanoplatelets (GNPs): the generator's plate geometry, the mid-plane triangulation, and the
assembly of tube-plate junctions, triangulation edges and plate-plate junctions against the
dictionary restatement of the reference (oracle/assembly.py).  No GPU."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla
from scipy.spatial import Delaunay

import oracle
from oracle import assembly
from cicy_geom_amd import network

ELEC = network.ElectricParams()


def small_params(seed=11):
    # 1.5 um RVE at the tube density of the 5 um / 40k-tube workload and the plate
    # density of the 8 um / 2k-plate one
    return network.NetworkParams(box=1.5, n_cnt=1080, n_gnp=14, seed=seed)


@pytest.fixture(scope="module")
def gnet():
    return network.Network(small_params())


def elec_dict(e):
    return dict(applied_voltage=e.applied_voltage, rho=e.resistivity_cnt, c1=e.c1, c2=e.c2,
                junction_exponential=e.junction_exponential,
                junction_resistance=e.junction_resistance,
                rho_gnp_t=e.resistivity_gnp_t, rho_gnp_surf=e.resistivity_gnp_surf)


def gnp_inputs(net, k):
    G = net.count("gnps")
    return dict(
        boundary_pts=[net.gnp_boundary_points(b) for b in range(6)],
        boundary_gnps=[net.boundary_gnps(b) for b in range(6)],
        cluster_gnps=[int(g) for g in net.cluster_gnps(k)],
        structure={g: net.gnp_structure(g) for g in range(G)},
        xyz=net.gnp_xyz,
        thick={g: net.gnp(g)[3] for g in range(G)},
        triangulation={g: [tuple(e) for e in net.gnp_triangulation(g)] for g in range(G)},
        mixed=[net.mixed_junction(j) for j in range(net.count("mixed_junctions"))],
        cluster_mixed=[int(j) for j in net.cluster_mixed_junctions(k)],
        gjunc=[net.gnp_junction(j) for j in range(net.count("gnp_junctions"))],
        cluster_gjunc=[int(j) for j in net.cluster_gnp_junctions(k)])


def restated(net, k, r_flag, family, elec=ELEC):
    tubes = [int(t) for t in net.cluster_tubes(k)]
    g = gnp_inputs(net, k)
    ref = assembly.assemble(family, r_flag, [net.boundary_points(b) for b in range(6)], tubes,
                            {t: net.elements(t) for t in tubes}, net.xyz, net.radii,
                            [net.junction(j) for j in range(net.count("junctions"))],
                            [int(j) for j in net.cluster_junctions(k)], elec_dict(elec),
                            net.params.box, gnp=g)
    return ref, g


def largest(net, direction=0):
    ks = net.percolating_clusters(direction)
    assert ks, "the test network must percolate"
    return max(ks, key=lambda c: net.cluster(c)[0])


def test_network_has_plates_in_the_percolating_cluster(gnet):
    k = largest(gnet)
    assert len(gnet.cluster_gnps(k)) >= 5
    assert len(gnet.cluster_mixed_junctions(k)) > 100
    s = gnet.assemble(k, 1, 0)
    assert s.gnp_nodes > 100


@pytest.mark.parametrize("r_flag,family", [(1, 0), (0, 0), (1, 1), (1, 2), (1, 6), (0, 6),
                                           (1, 3)])
def test_gnp_assembly_bit_identical_to_restatement(gnet, r_flag, family):
    k = largest(gnet)
    s = gnet.assemble(k, r_flag, family)
    ref, _ = restated(gnet, k, r_flag, family)
    assert s.nodes == ref["nodes"] and s.reserved == ref["reserved"]
    for name in ("row_ptr", "col_ind", "values", "diag", "b", "vef", "x0"):
        assert np.array_equal(getattr(s, name), ref[name]), name
    assert {p: int(n) for p, n in enumerate(s.lmm_gnp) if n >= 0} == ref["lmm_gnp"]


def test_gnp_boundary_elements_match_restatement(gnet):
    for d in range(3):
        ks = gnet.percolating_clusters(d)
        if not ks:
            continue
        k = max(ks, key=lambda c: gnet.cluster(c)[0])
        s = gnet.assemble(k, 1, d, direction=d)
        ref, g = restated(gnet, k, 1, d)
        lmm = {p: int(n) for p, n in enumerate(s.lmm) if n >= 0}
        out = assembly.boundary_elements(
            d, {b: gnet.boundary_tubes(b) for b in range(6)},
            {t: gnet.elements(t) for t in range(gnet.count("tubes"))}, lmm, gnet.xyz,
            gnet.radii, ELEC.resistivity_cnt, gnp=g, lmm_gnp=ref["lmm_gnp"],
            points_cnt_rad=ref["points_cnt_rad"], elec=elec_dict(ELEC))
        n_plate_edges = 0
        for (a, b, c), (x, y, z) in zip(s.boundary_elements, out):
            assert np.array_equal(a, x) and np.array_equal(b, y) and np.array_equal(c, z)
            n_plate_edges += int(np.sum(c < 1.0))  # plate edges are far below 1 Ohm
        assert n_plate_edges > 0


def test_current_conservation_with_plates(gnet):
    """Solved exactly, the currents through the two electrodes agree: every element at a
    boundary, tube or plate edge, is in the lists (the reference's own check,
    Electrical_analysis.cpp:396-402)."""
    k = largest(gnet)
    s = gnet.assemble(k, 1, 0, direction=0)
    n = s.n
    rows = np.repeat(np.arange(n), np.diff(s.row_ptr))
    L = sp.csr_matrix((s.values, (rows, s.col_ind)), shape=(n, n))
    A = (L + L.T + sp.diags(s.diag)).tocsc()
    x = sla.spsolve(A, s.b)
    v = np.concatenate([s.vef, x])
    i1, i2 = (oracle.boundary_current(*s.boundary_elements[w], v) for w in range(2))
    assert abs(i1 + i2) <= 1e-7 * abs(i1)
    # a planted mistake: dropping the plate edges from one list breaks the balance
    nb, ni, re = s.boundary_elements[0]
    keep = re >= 1.0
    j1 = oracle.boundary_current(nb[keep], ni[keep], re[keep], v)
    assert abs(j1 + i2) > 1e-4 * abs(i1)


def test_triangulation_is_delaunay_with_super_triangle(gnet):
    """Bowyer-Watson with the reference's super triangle yields the Delaunay triangulation
    of the points and the three super vertices; its edges between points must equal
    qhull's (scipy) for points in general position."""
    checked = 0
    for g in range(gnet.count("gnps")):
        st = gnet.gnp_structure(g)
        tri = gnet.gnp_triangulation(g)
        if len(tri) == 0:
            continue
        c, ax, l, _ = gnet.gnp(g)
        uv = (gnet.gnp_xyz[st] - c) @ ax[:2].T
        if len(st) <= 4:
            want = {frozenset((int(st[i]), int(st[j]))) for i in range(len(st))
                    for j in range(i + 1, len(st))}
        else:
            s3 = math.sqrt(3.0)
            sup = np.array([[0.0, (0.75 * s3 + 0.5) * l], [(0.75 + 0.5 * s3) * l, -l],
                            [-(0.75 + 0.5 * s3) * l, -l]])
            d = Delaunay(np.vstack([uv, sup]))
            want = set()
            for simp in d.simplices:
                for a, b in ((0, 1), (1, 2), (2, 0)):
                    i, j = int(simp[a]), int(simp[b])
                    if i < len(st) and j < len(st):
                        want.add(frozenset((int(st[i]), int(st[j]))))
        got = [frozenset((int(a), int(b))) for a, b in tri]
        assert len(got) == len(set(got)), "repeated edge"
        assert set(got) == want, g
        checked += 1
    assert checked >= 5


def test_resistance_gnp_closed_forms():
    """Calculate_resistance_gnp: rho_s along global x and y, rho_t along z, their mean at
    45 degrees in the xz-plane; R = rho L / (pi r1 r2)."""
    e = dict(rho_gnp_surf=2.0, rho_gnp_t=7.0)
    r1, r2, L = 0.005, 0.0025, 0.3
    base = L / (math.pi * r1 * r2)
    for d, rho in (((1, 0, 0), 2.0), ((0, 1, 0), 2.0), ((0, 0, 1), 7.0),
                   ((1 / math.sqrt(2), 0, 1 / math.sqrt(2)), 4.5)):
        p2 = [L * d[0], L * d[1], L * d[2]]
        assert assembly.resistance_gnp([0, 0, 0], p2, r1, r2, e) == pytest.approx(rho * base,
                                                                                 rel=1e-14)
    with pytest.raises(ValueError):
        assembly.resistance_gnp([1, 1, 1], [1, 1, 1], r1, r2, e)


def test_mixed_junction_law_uses_the_smaller_of_diameter_and_thickness():
    e = dict(junction_exponential=True, c1=0.7, c2=14000.0)
    d = 0.001
    f = 0.7 * d * math.exp(14000.0 * d)
    # tube radius 5 nm (diameter 10 nm) against a 4 nm plate: area 4 nm squared
    assert assembly.junction_resistance(d, 0.005, 0.004, e, True, False) == pytest.approx(
        f / 0.004 ** 2, rel=1e-14)
    # against a 30 nm plate: the tube's diameter limits
    assert assembly.junction_resistance(d, 0.005, 0.03, e, True, False) == pytest.approx(
        f / 0.01 ** 2, rel=1e-14)


def test_plate_geometry(gnet):
    box = gnet.params.box
    gx = gnet.gnp_xyz
    G = gnet.count("gnps")
    assert G == gnet.params.n_gnp
    plates = [gnet.gnp(g) for g in range(G)]
    for c, ax, l, t in plates:
        assert np.allclose(ax @ ax.T, np.eye(3), atol=1e-12)
    # tube-plate junctions: the plate point is on the plate's surface, the gap is the
    # clamped distance to the tube's axis point minus the radius
    r = gnet.params.radius
    for j in range(gnet.count("mixed_junctions")):
        p_cnt, p_gnp, cnt, g, dist = gnet.mixed_junction(j)
        c, ax, l, t = plates[g]
        q = (gx[p_gnp] - c) @ ax.T
        h = np.array([l / 2, l / 2, t / 2])
        assert np.all(np.abs(q) <= h + 1e-12)
        assert np.any(np.abs(np.abs(q) - h) < 1e-12)
        gap = max(np.linalg.norm(gnet.xyz[p_cnt] - gx[p_gnp]) - r, gnet.params.d_vdw)
        assert dist == pytest.approx(gap, abs=1e-12)
        assert gnet.params.d_vdw <= dist <= gnet.params.d_tunnel + 1e-10
    # plate-plate junctions: both points on their plate's surface, gap within the cutoff
    for j in range(gnet.count("gnp_junctions")):
        p1, p2, g1, g2, dist = gnet.gnp_junction(j)
        assert g1 < g2 and gnet.params.d_vdw <= dist <= gnet.params.d_tunnel + 1e-10
        for p, g in ((p1, g1), (p2, g2)):
            c, ax, l, t = plates[g]
            q = (gx[p] - c) @ ax.T
            assert abs(abs(q[2]) - t / 2) < 1e-12
    # boundary points sit on their faces; pruned structures keep only relevant faces
    face_axis = {4: (0, 0.0), 2: (0, box), 5: (1, 0.0), 3: (1, box), 1: (2, 0.0), 0: (2, box)}
    on_face = {}
    for b, (a, val) in face_axis.items():
        for p in gnet.gnp_boundary_points(b):
            assert abs(gx[p, a] - val) < 1e-9
            on_face[int(p)] = b
    for k in range(gnet.count("clusters")):
        fam = gnet.cluster(k)[1]
        if fam < 0 or fam == 6:
            continue
        rel = set(assembly.boundaries_of_family(fam))
        for g in gnet.cluster_gnps(k):
            for p in gnet.gnp_structure(int(g)):
                assert on_face.get(int(p), next(iter(rel))) in rel


def test_tube_only_workloads_unchanged():
    """Without plates the generator draws exactly the tube stream it always drew."""
    a = network.generate(1)
    assert a.count("gnps") == 0 and a.count("gnp_points") == 0
    s = network.largest_percolating_system(a, 0)
    assert s.gnp_nodes == 0 and len(s.lmm_gnp) == 0


def _flat(ptr, count, dtype=np.int64):
    import ctypes
    if count == 0:
        return np.zeros(0, dtype)
    ct = ctypes.c_int64 if dtype == np.int64 else ctypes.c_double
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), (count,)).copy()


def test_flat_network_arrays_match_the_accessors(gnet):
    """cicy_net_arrays_get (the device assembly's input) flattens every list of lists in
    order: tubes' element points, plates' structure and triangulation, face lists, and each
    cluster's tubes, plates and junction lists, as the per-item accessors return them."""
    import ctypes
    from cicy_geom_amd.dea import JunctionRec, NetArrays
    a = NetArrays()
    assert gnet._lib.cicy_net_arrays_get(gnet._h, ctypes.addressof(a))
    assert a.n_points == gnet.count("points") and a.n_tubes == gnet.count("tubes")
    assert a.n_gnps == gnet.count("gnps") and a.n_clusters == gnet.count("clusters")
    assert np.array_equal(_flat(a.xyz, 3 * a.n_points, np.float64), gnet.xyz.reshape(-1))
    ep = _flat(a.elem_ptr, a.n_tubes + 1)
    pts = _flat(a.elem_pts, int(ep[-1]))
    for t in range(0, a.n_tubes, 37):
        assert np.array_equal(pts[ep[t]:ep[t + 1]], gnet.elements(t))
    junc = np.ctypeslib.as_array(ctypes.cast(a.junctions, ctypes.POINTER(JunctionRec)),
                                 (a.n_junctions,))
    for j in range(0, a.n_junctions, 11):
        p1, p2, n1, n2, d = gnet.junction(j)
        assert (junc[j]["p1"], junc[j]["p2"], junc[j]["e1"], junc[j]["e2"], junc[j]["dist"]) == \
            (p1, p2, n1, n2, d)
    mixed = np.ctypeslib.as_array(ctypes.cast(a.mixed, ctypes.POINTER(JunctionRec)),
                                  (a.n_mixed,))
    for j in range(a.n_mixed):
        assert tuple(mixed[j].tolist()) == tuple(gnet.mixed_junction(j))
    gs_ptr = _flat(a.gstruct_ptr, a.n_gnps + 1)
    gs = _flat(a.gstruct, int(gs_ptr[-1]))
    gt_ptr = _flat(a.gtri_ptr, a.n_gnps + 1)
    gt = _flat(a.gtri, int(gt_ptr[-1]))
    thick = _flat(a.gnp_thickness, a.n_gnps, np.float64)
    for g in range(a.n_gnps):
        assert np.array_equal(gs[gs_ptr[g]:gs_ptr[g + 1]], gnet.gnp_structure(g))
        assert np.array_equal(gt[gt_ptr[g]:gt_ptr[g + 1]],
                              np.asarray(gnet.gnp_triangulation(g)).reshape(-1))
        assert thick[g] == gnet.gnp(g)[3]
    for name, fn in (("bpts", gnet.boundary_points), ("bcnt", gnet.boundary_tubes),
                     ("gbpts", gnet.gnp_boundary_points), ("gbnd", gnet.boundary_gnps)):
        fp = _flat(getattr(a, name + "_ptr"), 7)
        fe = _flat(getattr(a, name), int(fp[-1]))
        for b in range(6):
            assert np.array_equal(fe[fp[b]:fp[b + 1]], fn(b)), (name, b)
    for name, fn in (("cl_tubes", gnet.cluster_tubes), ("cl_junc", gnet.cluster_junctions),
                     ("cl_gnps", gnet.cluster_gnps), ("cl_mixed", gnet.cluster_mixed_junctions),
                     ("cl_gjunc", gnet.cluster_gnp_junctions)):
        cp = _flat(getattr(a, name + "_ptr"), a.n_clusters + 1)
        ce = _flat(getattr(a, name), int(cp[-1]))
        for k in range(a.n_clusters):
            assert np.array_equal(ce[cp[k]:cp[k + 1]], fn(k)), (name, k)
    assert list(a.window_lo) == [0.0] * 3 and list(a.window_hi) == [gnet.params.box] * 3
