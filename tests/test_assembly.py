"""The C++ generator/assembler (libcicy_net.so) against the dictionary-based restatement of
the reference's assembly (oracle/assembly.py), plus generator properties.  No GPU."""
import hashlib
import os

import numpy as np
import pytest

import oracle
from oracle import assembly
from cicy_geom_amd import network

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def net1():
    return network.generate(1)


def restated(net, k, r_flag, family, elec):
    tubes = [int(t) for t in net.cluster_tubes(k)]
    elements = {t: net.elements(t) for t in tubes}
    junctions = [net.junction(j) for j in range(net.count("junctions"))]
    bpts = [net.boundary_points(b) for b in range(6)]
    xyz = net.xyz
    return assembly.assemble(family, r_flag, bpts, tubes, elements, xyz, net.radii, junctions,
                             [int(j) for j in net.cluster_junctions(k)],
                             dict(applied_voltage=elec.applied_voltage, rho=elec.resistivity_cnt,
                                  c1=elec.c1, c2=elec.c2,
                                  junction_exponential=elec.junction_exponential,
                                  junction_resistance=elec.junction_resistance),
                             net.params.box)


@pytest.mark.parametrize("r_flag,family", [(1, 0), (0, 0), (1, 1), (1, 2), (0, 6), (1, 6),
                                           (1, 3), (0, 4)])
def test_assembly_bit_identical_to_restatement(net1, r_flag, family):
    elec = network.ElectricParams()
    k = max(range(net1.count("clusters")), key=lambda c: net1.cluster(c)[0])
    s = net1.assemble(k, r_flag, family, elec, direction=family if family < 3 else 0)
    ref = restated(net1, k, r_flag, family, elec)
    assert s.nodes == ref["nodes"] and s.reserved == ref["reserved"]
    for name in ("row_ptr", "col_ind", "values", "diag", "b", "vef", "x0"):
        assert np.array_equal(getattr(s, name), ref[name]), name
    lmm = {p: int(n) for p, n in enumerate(s.lmm) if n >= 0}
    assert lmm == ref["lmm"]
    if s.has_guess:
        r0 = oracle.update_residual(s.row_ptr, s.col_ind, s.values, s.diag, s.x0, s.b)
        assert np.array_equal(r0, s.r0)
    else:
        assert np.array_equal(s.r0, s.b) and not np.any(s.x0)


def test_boundary_elements_match_restatement(net1):
    elec = network.ElectricParams()
    for d in range(3):
        ks = net1.percolating_clusters(d)
        if not ks:
            continue
        s = net1.assemble(ks[0], 1, d, elec, direction=d)
        lmm = {p: int(n) for p, n in enumerate(s.lmm) if n >= 0}
        tubes = {b: net1.boundary_tubes(b) for b in range(6)}
        elements = {t: net1.elements(t) for t in range(net1.count("tubes"))}
        ref = assembly.boundary_elements(d, tubes, elements, lmm, net1.xyz, net1.radii,
                                         elec.resistivity_cnt)
        for (a, b, c), (x, y, z) in zip(s.boundary_elements, ref):
            assert np.array_equal(a, x) and np.array_equal(b, y) and np.array_equal(c, z)


def test_constant_junction_law(net1):
    elec = network.ElectricParams(junction_exponential=False, junction_resistance=1e7)
    k = net1.percolating_clusters(0)[0]
    s = net1.assemble(k, 1, 0, elec)
    ref = restated(net1, k, 1, 0, elec)
    assert np.array_equal(s.values, ref["values"]) and np.array_equal(s.diag, ref["diag"])


def test_c1_golden_solution():
    """Sparse direct (splu) solution of the 1 um workload pins the oracle's solve."""
    f = np.load(os.path.join(GOLDEN, "c1_solution.npz"), allow_pickle=False)
    s = network.largest_percolating_system(network.generate(1), 0)
    h = hashlib.sha256()
    for k in ("row_ptr", "col_ind", "values", "diag", "b"):
        h.update(np.ascontiguousarray(getattr(s, k)).tobytes())
    assert h.hexdigest() == str(f["sha256"]), "generator or assembly output changed"
    res = oracle.pcg_sss(s.nodes, s.reserved, 1e-10, s.row_ptr, s.col_ind, s.values, s.diag,
                         s.r0, s.vef, s.x0)
    err = np.linalg.norm(res["x"] - f["x"]) / np.linalg.norm(f["x"])
    assert res["status"] == oracle.ORC_OK and err < 1e-8
    xt, _, _ = oracle.cg_truth(s.row_ptr, s.col_ind, s.values, s.diag, s.b, tol=1e-14)
    assert np.linalg.norm(xt - f["x"]) / np.linalg.norm(f["x"]) < 1e-10


def test_generator_is_deterministic_and_sane():
    a = network.generate(1)
    b = network.generate(1)
    assert np.array_equal(a.xyz, b.xyz)
    assert a.count("junctions") == b.count("junctions")
    box = a.params.box
    xyz = a.xyz
    assert xyz.min() >= 0.0 and xyz.max() <= box
    # clipped ends sit exactly on their faces
    for bface, (ax, val) in {4: (0, 0.0), 2: (0, box), 5: (1, 0.0), 3: (1, box),
                             1: (2, 0.0), 0: (2, box)}.items():
        pts = a.boundary_points(bface)
        assert np.all(np.abs(xyz[pts, ax] - val) < 1e-10)
    # junction gaps respect the clamp and the cutoff
    for j in range(a.count("junctions")):
        _, _, n1, n2, d = a.junction(j)
        assert n1 < n2 and a.params.d_vdw <= d <= a.params.d_tunnel
    # segments are close to the nominal 20 nm
    cp = a.cnt_ptr
    seg = np.linalg.norm(np.diff(xyz, axis=0), axis=1)
    inner = np.ones(len(seg), bool)
    inner[cp[1:-1] - 1] = False
    assert seg[inner].max() < 0.0301


def test_config_sizes():
    s = network.largest_percolating_system(network.generate(1), 0)
    assert 3_000 < s.n < 9_000 and 1.0 < s.nnz / s.n < 1.2


def test_invalid_inputs(net1):
    with pytest.raises(RuntimeError):
        net1.assemble(0, 2, 0)
    with pytest.raises(RuntimeError):
        net1.assemble(0, 1, 7)
    with pytest.raises(ValueError):
        network.Network(network.NetworkParams(box=-1.0))
