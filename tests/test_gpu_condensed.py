"""GPU parity of the register-resident CG and of the condensed solve (through the C-ABI)
against the CPU oracle.

Bars: ||b - A x|| <= 1e-10 ||b|| checked with the oracle's SSS product, and the voltages
no further from the FP64 truth (a CG iterate at 1e-13) than the reference algorithm's own
solve at the same stopping rule, within a factor 3.  These networks span seven decades of
conductance: a 1e-10 residual leaves 1e-8 .. 3e-6 of relative error for any CG, the
reference's included, so an absolute error bar would test the conditioning, not the solver.
"""
import numpy as np
import pytest

import oracle
from cicy_geom_amd import dea, network

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dea.Context(0)
    yield c
    c.close()


def networks():
    out = {"c1": network.largest_percolating_system(network.generate(1), 0)}
    p = network.config_params(4, 0, 6000)
    net = network.Network(p)
    ks = net.percolating_clusters(0)
    c = max(ks, key=lambda q: net.cluster(q)[0])
    out["rve6k"] = net.assemble(c, 1, 0, direction=0)
    net.close()
    return out


@pytest.fixture(scope="module")
def systems():
    return networks()


@pytest.fixture(scope="module")
def truths(systems):
    return {k: oracle.cg_truth(s.row_ptr, s.col_ind, s.values, s.diag, s.b, tol=1e-13)[0]
            for k, s in systems.items()}


def check(s, x, truth, err_ref, tol=1e-10):
    ax = oracle.spmv_sss(s.row_ptr, s.col_ind, s.values, s.diag, x)
    tr = np.linalg.norm(s.b - ax) / np.linalg.norm(s.b)
    err = np.linalg.norm(x - truth) / np.linalg.norm(truth)
    assert tr <= tol * (1 + 1e-6), tr
    assert err <= max(3.0 * err_ref, 1e-9), (err, err_ref)
    return tr, err


@pytest.fixture(scope="module")
def ref_errors(ctx, systems, truths):
    """Error of the reference algorithm (FP64 Jacobi-PCG) at ||b - Ax|| <= 1e-10 ||b||."""
    out = {}
    for k, s in systems.items():
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        sysd.set_rhs(s.b, None, s.b)
        st = sysd.solve(dea.METHOD_PCG_F64, 1e-10, 0, dea.REL_B)
        x, _ = sysd.get()
        sysd.close()
        assert st.status == dea.CICY_OK
        out[k] = np.linalg.norm(x - truths[k]) / np.linalg.norm(truths[k])
        print("reference-algorithm error", k, out[k], st.iters, flush=True)
    return out


@pytest.mark.parametrize("name", ["c1", "rve6k"])
@pytest.mark.parametrize("method", [dea.METHOD_RESIDENT, dea.METHOD_CONDENSED,
                                    dea.METHOD_CONDENSED_RESIDENT])
def test_solution_and_true_residual(ctx, systems, truths, ref_errors, name, method):
    s = systems[name]
    sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    try:
        sysd.set_rhs(s.b, None, s.b)
        st = sysd.solve(method, 1e-10, 0, dea.REL_B)
        x, _ = sysd.get()
        print(name, dea.METHOD_NAMES[method], st.as_dict(), sysd.size("resident_rows"),
              sysd.size("condensed_rows"), flush=True)
        assert st.status == dea.CICY_OK
        tr, err = check(s, x, truths[name], ref_errors[name])
        # both are ||b - A x|| / ||b|| in FP64, summed in different orders: they agree to
        # 1e-3 relative, or to rounding (a few 1e-16) when the solve reaches machine precision
        assert abs(st.true_rel_res - tr) <= 1e-3 * tr + 2e-15
        if method == dea.METHOD_RESIDENT:
            # the same CG as the two-kernel solve: iteration counts agree closely
            ref = sysd.solve(dea.METHOD_PCG_F64, 1e-10, 0, dea.REL_B)
            assert abs(ref.iters - st.iters) <= max(16, ref.iters // 10)
        else:
            assert 0 < sysd.size("condensed_rows") < s.n // 2
    finally:
        sysd.close()


@pytest.mark.parametrize("rows", [4, 8])
def test_resident_forced_rows_several_groups(ctx, systems, truths, ref_errors, rows):
    """Forced rows per thread spread the 6k-tube system over many work-groups."""
    s = systems["rve6k"]
    ctx.set_resident_rows(rows)
    try:
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    finally:
        ctx.set_resident_rows(0)
    try:
        sysd.prepare(dea.METHOD_RESIDENT)  # built with the rows asked for at creation
        assert sysd.size("resident_rows") == rows
        assert sysd.size("resident_groups") == -(-s.n // (512 * rows))
        sysd.set_rhs(s.b, None, s.b)
        st = sysd.solve(dea.METHOD_RESIDENT, 1e-10, 0, dea.REL_B)
        x, _ = sysd.get()
        assert st.status == dea.CICY_OK
        check(s, x, truths["rve6k"], ref_errors["rve6k"])
        # a second solve on the same data repeats the first (state is reset per solve)
        st2 = sysd.solve(dea.METHOD_RESIDENT, 1e-10, 0, dea.REL_B)
        x2, _ = sysd.get()
        assert st2.status == dea.CICY_OK
        check(s, x2, truths["rve6k"], ref_errors["rve6k"])
    finally:
        sysd.close()


def test_resident_iteration_cap_and_chunks(ctx, systems):
    """max_iter beyond one cooperative launch (2048 iterations) and below convergence."""
    s = systems["rve6k"]
    sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    try:
        sysd.set_rhs(s.b, None, s.b)
        st = sysd.solve(dea.METHOD_RESIDENT, 1e-30, 3000, dea.REL_B)
        assert st.status == dea.CICY_MAXITER and st.iters == 3000
    finally:
        sysd.close()


@pytest.mark.parametrize("method", [dea.METHOD_CONDENSED, dea.METHOD_CONDENSED_RESIDENT])
def test_condensed_reference_rule_with_ramp_guess(ctx, systems, method):
    """CICY_REL_R0 from the reference's ramp guess: ||b - A x|| <= tol ||r0||."""
    s = systems["rve6k"]
    assert s.has_guess
    sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    try:
        sysd.set_rhs(s.r0, s.x0, s.b)
        st = sysd.solve(method, 1e-8, 0, dea.REL_R0)
        x, _ = sysd.get()
        assert st.status == dea.CICY_OK
        ax = oracle.spmv_sss(s.row_ptr, s.col_ind, s.values, s.diag, x)
        assert np.linalg.norm(s.b - ax) <= 1e-8 * np.linalg.norm(s.r0) * (1 + 1e-6)
    finally:
        sysd.close()


def test_condensed_zero_rhs_and_random_laplacian(ctx):
    """Zero right-hand side returns zero at once; an unstructured matrix has no condensed
    system and the method says so."""
    s = network.largest_percolating_system(network.generate(1), 0)
    sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    try:
        z = np.zeros(s.n)
        sysd.set_rhs(z, None, None)
        st = sysd.solve(dea.METHOD_CONDENSED, 1e-10, 0, dea.REL_R0)
        x, _ = sysd.get()
        assert st.status == dea.CICY_OK and not np.any(x)
    finally:
        sysd.close()
    rng = np.random.default_rng(5)
    n = 400
    pairs = {}
    for a, b in rng.integers(0, n, (2000, 2)):
        if abs(a - b) > 1:
            pairs[(max(a, b), min(a, b))] = -1.0
    keys = sorted(pairs)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int64)
    sysd = ctx.system(rp, col, -np.ones(len(col)), np.full(n, 30.0))
    try:
        assert sysd.size("condensed_rows") == 0
        sysd.set_rhs(np.ones(n), None, np.ones(n))
        with pytest.raises(dea.DEAError):
            sysd.solve(dea.METHOD_CONDENSED, 1e-10, 0, dea.REL_B)
    finally:
        sysd.close()


def test_dropin_call_with_condensed_method(ctx, systems):
    """Solve_DEA_equations_CG_SSS's call, solved by the condensed method: the reference's
    stopping rule (||r|| <= tol ||r0||) holds for the true residual, R comes back as
    b - A x and voltages = [VEF | x]."""
    s = systems["rve6k"]
    ctx.set_dropin_method(dea.METHOD_CONDENSED)
    try:
        ok, volt, x, R, st = ctx.solve_dea_cg_sss(s.nodes, s.reserved, 1e-9, s.col_ind,
                                                  s.row_ptr, s.values, s.diag, s.r0, s.vef,
                                                  s.x0)
    finally:
        ctx.set_dropin_method(dea.METHOD_PCG_F64)
    assert ok and st.status == dea.CICY_OK
    ax = oracle.spmv_sss(s.row_ptr, s.col_ind, s.values, s.diag, x)
    r_true = s.b - ax
    assert np.linalg.norm(r_true) <= 1e-9 * np.linalg.norm(s.r0) * (1 + 1e-6)
    # R is r0 - A d on the device, r_true = b - A (x0 + d) on the host: equal up to the
    # rounding of the two ways of forming it
    assert np.linalg.norm(R - r_true) <= 0.05 * np.linalg.norm(r_true) + 1e-12 * np.linalg.norm(s.b)
    assert np.array_equal(volt[:s.reserved], s.vef) and np.array_equal(volt[s.reserved:], x)


def test_prepare_builds_layouts_on_demand(ctx, systems):
    """cicy_sys_create uploads only the two-kernel layout; cicy_sys_prepare builds a
    method's layouts once (idempotent), and a solve prepares by itself."""
    s = systems["rve6k"]
    sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    try:
        assert sysd.size("condensed_rows") == 0 and sysd.size("core_blocks") == 0
        bytes0 = sysd.size("bytes")
        sysd.prepare(dea.METHOD_CONDENSED_BLOCK)
        m, core = sysd.size("condensed_rows"), sysd.size("core_rows")
        assert 0 < core <= m < s.n // 2
        assert sysd.size("core_groups") > 0
        sysd.prepare(dea.METHOD_CONDENSED_BLOCK)  # nothing rebuilt
        assert (sysd.size("condensed_rows"), sysd.size("core_rows")) == (m, core)
        sysd.set_rhs(s.b, None, s.b)
        st = sysd.solve(dea.METHOD_CONDENSED_BLOCK, 1e-10, 0, dea.REL_B)
        assert st.status == dea.CICY_OK and st.true_rel_res <= 1e-10 * (1 + 1e-6)
        assert sysd.size("bytes") == bytes0  # no FP32 copy yet
        st = sysd.solve(dea.METHOD_MIXED, 1e-6, 2000, dea.REL_B)  # built by this solve
        assert st.status in (dea.CICY_OK, dea.CICY_MAXITER)
        assert sysd.size("bytes") > bytes0
        with pytest.raises(dea.DEAError):
            sysd.prepare(99)
    finally:
        sysd.close()
