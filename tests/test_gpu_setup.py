"""GPU parity of the device set-up (cicy_sys_create_device and the device condensation,
series rounds and core layout) against the host forms of the same steps.

The host forms (build_band, build_condensed_host, build_series_host, build_core_host) are
what the CPU tests check against dense Schur complements, the series law and direct solves.
The device set-up must give the same system bit for bit, so every solve on it is the same
solve: the same iterations, and the same x to the last bit.  Checked for the plain FP64
PCG on the band layout built from device arrays, and for the condensed methods on networks
with and without nanoplatelets, up to the full 8 um RVE with 160k tubes and 2k plates."""
import numpy as np
import pytest

from cicy_geom_amd import dea, network

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dea.Context(0)
    yield c
    c.close()


def plate_case():
    p = network.NetworkParams(box=2.0, n_cnt=2560, n_gnp=31, seed=5)
    net = network.Network(p)
    k = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    s = net.assemble(k, 1, 0, direction=0)
    ranges = network.plate_row_ranges(net, k, s)
    net.close()
    return s, ranges


@pytest.fixture(scope="module")
def cases():
    out = {"plates": plate_case()}
    out["c1"] = (network.largest_percolating_system(network.generate(1), 0),
                 np.zeros((0, 2), np.int64))
    net = network.Network(network.config_params(4, 0, 6000))
    k = max(net.percolating_clusters(0), key=lambda q: net.cluster(q)[0])
    out["rve6k"] = (net.assemble(k, 1, 0, direction=0), np.zeros((0, 2), np.int64))
    net.close()
    return out


SIZES = ("condensed_rows", "condensed_nnz", "core_rows", "core_nnz", "series_rounds",
         "core_blocks", "core_block_rows", "core_block_entries", "core_groups")


def solve(ctx, s, ranges, method, where, from_device=False, max_iter=0, tol=1e-10):
    """One system set up `where` ("host" / "device"), solved with `method`; returns the
    stats, x and the layout sizes."""
    ctx.set_setup(where)
    keep = []
    try:
        if from_device:
            arrs = [ctx.upload(np.ascontiguousarray(a, dt)) for a, dt in
                    ((s.row_ptr, np.int64), (s.col_ind, np.int64), (s.values, np.float64),
                     (s.diag, np.float64))]
            keep += arrs
            sysd = ctx.system_device(s.n, len(s.values), *arrs)
        else:
            sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        sysd.set_rhs(s.b, None, s.b)
        if method == dea.METHOD_CONDENSED_BLOCK:
            sysd.set_blocks(ranges)
        st = sysd.solve(method, tol, max_iter, dea.REL_B)
        x, _ = sysd.get()
        info = {k: sysd.size(k) for k in SIZES}
        info["format"] = sysd.size("format")
        info["off_layout"] = sysd.size("off_layout")
        sysd.close()
    finally:
        ctx.set_setup("device")
        for a in keep:
            a.close()
    return st, x, info


@pytest.mark.parametrize("name", ["plates", "c1", "rve6k"])
@pytest.mark.parametrize("method", [dea.METHOD_CONDENSED_BLOCK, dea.METHOD_CONDENSED])
def test_device_setup_equals_host_setup(ctx, cases, name, method):
    s, ranges = cases[name]
    st_h, x_h, info_h = solve(ctx, s, ranges, method, "host")
    st_d, x_d, info_d = solve(ctx, s, ranges, method, "device")
    print(f"\n{name} {dea.METHOD_NAMES[method]}: host {info_h}\n   device {info_d}; iters "
          f"{st_h.iters} / {st_d.iters}")
    assert info_d == info_h
    assert info_d["condensed_rows"] > 0 and info_d["core_rows"] > 0
    assert st_h.status == dea.CICY_OK and st_d.status == dea.CICY_OK
    assert st_d.iters == st_h.iters and st_d.refinements == st_h.refinements
    assert np.array_equal(x_d, x_h)


@pytest.mark.parametrize("name", ["plates", "c1", "rve6k"])
def test_device_arrays_give_the_host_layout(ctx, cases, name):
    """cicy_sys_create_device from the reference's int64 arrays lays the system out as
    cicy_sys_create does: the same band + remainder layout, hence the same plain FP64 PCG
    iterations bit for bit (a 300-iteration prefix and a full solve)."""
    s, ranges = cases[name]
    for max_iter in (300, 0):
        st_h, x_h, info_h = solve(ctx, s, ranges, dea.METHOD_PCG_F64, "device", max_iter=max_iter)
        st_d, x_d, info_d = solve(ctx, s, ranges, dea.METHOD_PCG_F64, "device", from_device=True,
                                  max_iter=max_iter)
        assert info_d["format"] == info_h["format"] == dea.FORMAT_BAND
        assert info_d["off_layout"] == info_h["off_layout"]
        assert st_d.iters == st_h.iters
        assert np.array_equal(x_d, x_h)
    st_h, x_h, _ = solve(ctx, s, ranges, dea.METHOD_CONDENSED_BLOCK, "host")
    st_d, x_d, _ = solve(ctx, s, ranges, dea.METHOD_CONDENSED_BLOCK, "device", from_device=True)
    assert st_d.iters == st_h.iters and np.array_equal(x_d, x_h)


def test_device_rhs_and_device_arrays_end_to_end(ctx, cases):
    """Everything from device memory: system, r0 and b; the reference norms summed on the
    device (fixed order) agree with the host sums to rounding, and the solve converges to the
    same residual bar."""
    s, ranges = cases["plates"]
    arrs = [ctx.upload(np.ascontiguousarray(a, dt)) for a, dt in
            ((s.row_ptr, np.int64), (s.col_ind, np.int64), (s.values, np.float64),
             (s.diag, np.float64), (s.b, np.float64))]
    try:
        sysd = ctx.system_device(s.n, len(s.values), *arrs[:4])
        sysd.set_rhs_device(arrs[4], None, arrs[4])
        sysd.set_blocks(ranges)
        st = sysd.solve(dea.METHOD_CONDENSED_BLOCK, 1e-10, 0, dea.REL_B)
        x, _ = sysd.get()
        sysd.close()
    finally:
        for a in arrs:
            a.close()
    st_h, x_h, _ = solve(ctx, s, ranges, dea.METHOD_CONDENSED_BLOCK, "device")
    assert st.status == dea.CICY_OK
    rows = np.repeat(np.arange(s.n), np.diff(s.row_ptr))
    ax = s.diag * x
    np.add.at(ax, rows, s.values * x[s.col_ind])
    np.add.at(ax, s.col_ind, s.values * x[rows])
    assert np.linalg.norm(s.b - ax) <= 1e-10 * np.linalg.norm(s.b) * (1 + 1e-9)
    assert np.linalg.norm(x - x_h) <= 1e-9 * np.linalg.norm(x_h)


def test_device_arrays_refused_when_malformed(ctx, cases):
    """Malformed device arrays are refused with the first bad row, as on the host; the checks
    read a row's entries only inside its checked bounds."""
    s, _ = cases["c1"]
    n, nnz = s.n, len(s.values)

    def attempt(rp, ci, dg):
        arrs = [ctx.upload(np.ascontiguousarray(a, dt)) for a, dt in
                ((rp, np.int64), (ci, np.int64), (s.values, np.float64), (dg, np.float64))]
        try:
            with pytest.raises(dea.DEAError) as e:
                ctx.system_device(n, nnz, *arrs)
            return str(e.value)
        finally:
            for a in arrs:
                a.close()

    r = int(np.argmax(np.diff(s.row_ptr) > 0))  # first row with an entry
    ci = np.array(s.col_ind)
    ci[s.row_ptr[r]] = r  # not strictly below its row
    assert f"is not strictly below row {r}" in attempt(s.row_ptr, ci, s.diag)
    rp = np.array(s.row_ptr)
    rp[n // 2] = rp[n // 2 - 1] - 1  # row n/2 - 1 ends before it starts
    assert rp[n // 2] >= 0
    assert f"row_ptr decreases or exceeds nnz at row {n // 2 - 1}" in attempt(rp, s.col_ind, s.diag)
    dg = np.array(s.diag)
    dg[7] = 0.0
    assert "zero or non-finite diagonal at row 7" in attempt(s.row_ptr, s.col_ind, dg)
    rp = np.array(s.row_ptr)
    rp[-1] = nnz - 1
    assert "row_ptr must start at 0 and end at nnz" in attempt(rp, s.col_ind, s.diag)
    # mixed and resident layouts need host arrays
    arrs = [ctx.upload(np.ascontiguousarray(a, dt)) for a, dt in
            ((s.row_ptr, np.int64), (s.col_ind, np.int64), (s.values, np.float64),
             (s.diag, np.float64))]
    try:
        sysd = ctx.system_device(n, nnz, *arrs)
        sysd.set_rhs(s.b, None, s.b)
        with pytest.raises(dea.DEAError, match="needs the host arrays"):
            sysd.solve(dea.METHOD_MIXED, 1e-10, 0, dea.REL_B)
        sysd.close()
    finally:
        for a in arrs:
            a.close()


def test_device_setup_config3_equals_host_setup(ctx):
    """The benchmark's system (8 um RVE, 160k tubes + 2k plates, ~6M rows): device and host
    set-ups give the same condensation, core layout and block-method solve, bit for bit."""
    net = network.generate(3)
    k = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    s = net.assemble(k, 1, 0, direction=0)
    ranges = network.plate_row_ranges(net, k, s)
    del net
    st_h, x_h, info_h = solve(ctx, s, ranges, dea.METHOD_CONDENSED_BLOCK, "host")
    st_d, x_d, info_d = solve(ctx, s, ranges, dea.METHOD_CONDENSED_BLOCK, "device",
                              from_device=True)
    print(f"\nconfig 3: {info_d}; iters {st_d.iters}")
    assert info_d == {k: v for k, v in info_h.items()}
    assert st_d.status == dea.CICY_OK and st_d.iters == st_h.iters
    assert np.array_equal(x_d, x_h)


def test_floating_group_reduced_to_one_node_solves(ctx):
    """A floating group that the series rounds reduce to one node (diagonal 0, no
    neighbours) is decoupled with diagonal 1 by both set-ups, bit for bit, and the block
    method solves the system (it used to refuse it)."""
    from test_condensed_host import floating_triangle_system
    _, (rp, col, val, diag), n0 = floating_triangle_system()
    b = np.zeros(len(diag))
    b[0] = 1.0
    out = []
    for where in ("host", "device"):
        ctx.set_setup(where)
        try:
            arrs = [ctx.upload(np.ascontiguousarray(a, dt)) for a, dt in
                    ((rp, np.int64), (col, np.int64), (val, np.float64), (diag, np.float64))]
            sysd = ctx.system_device(len(diag), len(val), *arrs)
            sysd.set_rhs(b, None, b)
            st = sysd.solve(dea.METHOD_CONDENSED_BLOCK, 1e-12, 0, dea.REL_B)
            x, _ = sysd.get()
            out.append((st, x, sysd.size("core_rows")))
            sysd.close()
            for a in arrs:
                a.close()
        finally:
            ctx.set_setup("device")
    (st_h, x_h, core_h), (st_d, x_d, core_d) = out
    assert st_h.status == dea.CICY_OK and st_d.status == dea.CICY_OK
    assert core_h == core_d > 0
    assert np.array_equal(x_h, x_d)
    rows = np.repeat(np.arange(len(diag)), np.diff(rp))
    ax = diag * x_d
    np.add.at(ax, rows, val * x_d[col])
    np.add.at(ax, col, val * x_d[rows])
    assert np.linalg.norm(b - ax) <= 1e-12 * np.linalg.norm(b)


def test_configuration4_floating_component_solves(ctx):
    """Configuration 4, realisation 16 (seed 4 000 016): the x-direction system holds a
    floating group of tubes that the rounds reduce to one node.  Assembled on the device and
    solved by the block method it converges, and its resistance equals the plain FP64
    PCG's on the host-assembled system within 1e-9."""
    net = network.generate(4, realisation=16)
    k = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    s = net.assemble(k, 1, 0, direction=0)
    dn = net.device(ctx)
    a = dn.assemble(k, 1, 0, None, 0)
    sd = a.system()
    st = sd.solve(dea.METHOD_CONDENSED_BLOCK, 1e-12, 0, dea.REL_R0)
    x_b, _ = sd.get()
    sd.close()
    cur_b = ctx.boundary_current(*a.boundary_elements()[0], np.concatenate([a.vef, x_b]))
    a.close()
    dn.close()
    assert st.status == dea.CICY_OK, st.as_dict()
    sh = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    sh.set_rhs(s.r0, s.x0, s.b)
    st_p = sh.solve(dea.METHOD_PCG_F64, 1e-12, 0, dea.REL_R0)
    x_p, _ = sh.get()
    sh.close()
    cur_p = ctx.boundary_current(*s.boundary_elements[0], np.concatenate([s.vef, x_p]))
    print(f"\nconfig 4 realisation 16: N={s.n}, block {st.iters} core iterations, plain "
          f"{st_p.iters}; I {cur_b:.12e} / {cur_p:.12e}")
    assert st_p.status == dea.CICY_OK
    assert abs(cur_b - cur_p) <= 1e-9 * abs(cur_p)
    net.close()
