"""Resistance combination per direction (Calculate_resistances_and_resistivities,
Calculate_matrix_resistance, Calculate_resistivity; Electrical_analysis.cpp:873-987)
in cicy_geom_amd.direct_electrifying, for zero, one and many clusters.  No GPU: these
are host functions over the resistances the GPU solves produce."""
import math

import pytest

import oracle
from cicy_geom_amd import direct_electrifying as de
from cicy_geom_amd.network import ElectricParams


class Cuboid:
    """The reference's window members (Geometry_3D.h cuboid)."""
    def __init__(self, lx, wy, hz):
        self.len_x, self.wid_y, self.hei_z = lx, wy, hz


def test_matrix_resistance_closed_form():
    w = Cuboid(2.0, 3.0, 5.0)
    rho = 1.03e20
    assert de.calculate_matrix_resistance(0, w, rho) == rho * 2.0 / (3.0 * 5.0)
    assert de.calculate_matrix_resistance(1, w, rho) == rho * 3.0 / (2.0 * 5.0)
    assert de.calculate_matrix_resistance(2, w, rho) == rho * 5.0 / (2.0 * 3.0)
    # a cube of side a: rho / a along every axis
    for d in range(3):
        assert de.calculate_matrix_resistance(d, 8.0, rho) == rho * 8.0 / 64.0
    with pytest.raises(ValueError):
        de.calculate_matrix_resistance(3, w, rho)


def test_resistivity_inverts_resistance():
    w = Cuboid(2.0, 3.0, 5.0)
    for d, (length, area) in enumerate([(2.0, 15.0), (3.0, 10.0), (5.0, 6.0)]):
        assert de.calculate_resistivity(d, w, 7.0) == 7.0 * area / length
        # the matrix: resistance -> resistivity gives rho back
        rm = de.calculate_matrix_resistance(d, w, 4.0)
        assert math.isclose(de.calculate_resistivity(d, w, rm), 4.0, rel_tol=1e-15)


def test_zero_one_many_clusters():
    ep = ElectricParams(resistivity_matrix=1e20)
    w = Cuboid(1.0, 2.0, 4.0)
    par = [[], [250.0], [100.0, 300.0, 600.0]]
    R, rho = de.calculate_resistances_and_resistivities(w, ep, par)
    # x: no cluster -> the matrix (resistance and resistivity)
    assert R[0] == 1e20 * 1.0 / 8.0 and rho[0] == 1e20
    # y: one cluster -> that resistance
    assert R[1] == 250.0 and rho[1] == 250.0 * 4.0 / 2.0
    # z: three in parallel -> 1/(1/100 + 1/300 + 1/600) = 66.666...
    assert math.isclose(R[2], 200.0 / 3.0, rel_tol=1e-15)
    assert R[2] == oracle.parallel_resistance([100.0, 300.0, 600.0])
    assert rho[2] == R[2] * 2.0 / 4.0


def test_parallel_sum_order_matches_oracle():
    rk = [1.0e3 * (1 + 0.37 * k) for k in range(17)]
    assert de.calculate_resistance(rk, 0.0) == oracle.parallel_resistance(rk)
    assert de.calculate_resistance([], 5.0) == 5.0 == oracle.parallel_resistance([], 5.0)


def test_planted_inf_for_matrix_fails():
    """The earlier form returned inf for a direction without clusters; the reference takes
    the matrix resistance of the window."""
    ep = ElectricParams()
    R, _ = de.calculate_resistances_and_resistivities(1.0, ep, [[], [], []])
    assert all(math.isfinite(r) and r == ep.resistivity_matrix for r in R)
