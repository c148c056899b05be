"""The ramp initial guess (Initial_guess_for_CG / Initial_guess_for_cnt_nodes,
Direct_Electrifying.cpp:1396-1470; applied at :80-95) pinned by hand-computed values.

oracle.assembly.ramp_guess is checked on a three-tube network whose expected guess is
worked out by hand below, and each plausible mistake (min for max, the wrong axis, the
reserved nodes not skipped, the guess applied with R_flag = 0 or four reserved nodes)
is planted in a copy of the oracle function and shown to fail the same check.  The
product assembler (libcicy_net.so) is then held to the same closed form on the 1 um
workload, where every free node is a tube point.  No GPU."""
import numpy as np
import pytest

from oracle import assembly
from cicy_geom_amd import network

BOX = 4.0
V_APP = 2.0          # V_app / L_window = 0.5

# points (x, y, z); tube 0: 0-3, tube 1: 4-6, tube 2: 7-9
XYZ = np.array([
    [0.0, 1.0, 1.0], [1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [3.0, 1.0, 1.0],
    [2.5, 2.0, 1.0], [3.5, 2.0, 1.0], [4.0, 2.0, 1.0],
    [1.0, 3.0, 2.0], [2.0, 3.0, 2.0], [3.0, 0.5, 2.0],
])
TUBES = [2, 0, 1]          # cluster order: the tube holding point 6 comes last
ELEMENTS = {0: np.array([0, 1, 2, 3]), 1: np.array([4, 5, 6]), 2: np.array([7, 8, 9])}
RADII = np.full(3, 0.005)
# (P1, P2, tube1, tube2, distance)
JUNCTIONS = [(3, 4, 0, 1, 0.001), (1, 7, 0, 2, 0.001), (9, 5, 2, 1, 0.001)]


def boundary_points(family):
    """Point 0 on the first face of the family's axis, point 6 on the second."""
    b = [np.zeros(0, np.int64) for _ in range(6)]
    first, second = assembly.boundaries_of_family(family)[:2]
    b[first] = np.array([0])
    b[second] = np.array([6])
    return b


def build(family, r_flag, ramp=assembly.ramp_guess):
    elec = dict(applied_voltage=V_APP, rho=20.0, c1=0.692475, c2=14490.5,
                junction_exponential=True, junction_resistance=1e7)
    saved = assembly.ramp_guess
    assembly.ramp_guess = ramp
    try:
        return assembly.assemble(family, r_flag, boundary_points(family), TUBES, ELEMENTS,
                                 XYZ, RADII, JUNCTIONS, [0, 1, 2], elec, BOX)
    finally:
        assembly.ramp_guess = saved


# Node numbering (LM_matrix_for_cnts): point 0 -> node 0 and point 6 -> node 1 (boundary
# points); then the cluster's tubes in order: tube 2: points 7, 8, 9 -> nodes 2, 3, 4;
# tube 0: points 1, 2, 3 -> 5, 6, 7; tube 1: points 4, 5 -> 8, 9.  Free node i is row
# i - 2, so the rows hold points 7 8 9 1 2 3 4 5.
# x0[row] = 0.5 * (4 - coordinate of the row's point along the family's axis):
EXPECTED = {
    0: [1.5, 1.0, 0.5, 1.5, 1.0, 0.5, 0.75, 0.25],      # x
    1: [0.5, 0.5, 1.75, 1.5, 1.5, 1.5, 1.0, 1.0],       # y
    2: [1.0, 1.0, 1.0, 1.5, 1.5, 1.5, 1.5, 1.5],        # z
}


@pytest.mark.parametrize("family", [0, 1, 2])
def test_ramp_matches_hand_values(family):
    s = build(family, 1)
    assert s["reserved"] == 2 and s["nodes"] == 10
    assert s["lmm"] == {0: 0, 6: 1, 7: 2, 8: 3, 9: 4, 1: 5, 2: 6, 3: 7, 4: 8, 5: 9}
    assert np.array_equal(s["x0"], np.array(EXPECTED[family]))


def test_no_ramp_for_unit_resistors_or_four_reserved_nodes():
    assert not np.any(build(0, 0)["x0"])            # R_flag = 0: zero guess
    s = build(3, 1)                                 # family 3: x and y faces, R_d = 4
    assert s["reserved"] == 4 and not np.any(s["x0"])


# ---- planted mistakes: each mutant of the oracle function must fail the hand values

def _mutant(kind):
    def ramp(n, rd, r_flag, family, cluster_tubes, elements, lmm, xyz, applied_voltage, box):
        x0 = np.zeros(n)
        if kind == "any_flag" or (r_flag == 1 and rd == 2):
            vpc = applied_voltage / box
            ax = (family + 1) % 3 if kind == "axis" else family
            for t in cluster_tubes:
                for p in elements[t]:
                    nd = lmm[int(p)]
                    if kind == "no_skip" or nd >= 2:
                        if kind == "min":
                            v = vpc * (0.0 - xyz[int(p)][ax])
                        else:
                            v = vpc * (box - xyz[int(p)][ax])
                        x0[nd - rd] = v     # a reserved node wraps onto the last rows
        return x0
    return ramp


@pytest.mark.parametrize("kind", ["min", "axis", "no_skip"])
def test_planted_ramp_mistakes_fail_the_hand_values(kind):
    bad = build(0, 1, ramp=_mutant(kind))["x0"]
    assert not np.array_equal(bad, np.array(EXPECTED[0]))


def test_planted_flag_mistake_fails():
    assert np.any(build(0, 0, ramp=_mutant("any_flag"))["x0"])


def test_product_ramp_is_the_closed_form_on_the_1um_workload():
    """libcicy_net.so's x0 on configuration 1 (every free node a tube point) equals
    V_app / L * (max - coordinate) along each direction's axis, bit for bit."""
    net = network.generate(1)
    elec = network.ElectricParams()
    L = net.params.box
    for d in range(3):
        ks = net.percolating_clusters(d)
        if not ks:
            continue
        s = network.largest_percolating_system(net, d, with_xyz=True)
        assert s.reserved == 2 and s.has_guess
        want = (elec.applied_voltage / L) * (L - s.node_xyz[:, d])
        assert np.array_equal(s.x0, want), d
