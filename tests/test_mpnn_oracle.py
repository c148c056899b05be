"""Pins of the warm-start network's float64 oracle (oracle/mpnn.py): central finite
differences of its loss match its analytic gradient; a planted mistake in the backward
(a dropped gate derivative) is caught by the same check.  No GPU."""
import numpy as np

from oracle import mpnn


def small_graph(seed=0, n=40):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(1, n):
        rows.append(i); cols.append(i - 1); vals.append(-10 ** rng.uniform(-4, -2.5))
    for _ in range(n // 2):
        a, b = rng.integers(0, n, 2)
        if a != b and (max(a, b), min(a, b)) not in zip(rows, cols):
            rows.append(max(a, b)); cols.append(min(a, b)); vals.append(-10 ** rng.uniform(-10, -3))
    order = np.lexsort((cols, rows))
    rows = np.array(rows)[order]; cols = np.array(cols, np.int64)[order]; vals = np.array(vals)[order]
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    diag = np.zeros(n)
    np.add.at(diag, rows, -vals); np.add.at(diag, cols, -vals)
    diag[0] += 1e-3; diag[-1] += 1e-3
    b = np.zeros(n); b[0] = 1e-3
    xyz = rng.uniform(0, 2.0, (n, 3))
    gr = mpnn.graph_from_sss(row_ptr, cols, vals, diag, b, xyz, 2.0, 0, 1.0)
    vt = rng.uniform(0, 1, n)
    return gr, vt


def test_param_count():
    assert mpnn.n_params() == 32 * 4 + 32 + 3 * (2 * 32 * 32 + 3 * 32) + 32 + 1


def test_gradient_matches_finite_differences():
    gr, vt = small_graph()
    theta = mpnn.init_params(3)
    # make the correction non-trivial so every path carries gradient
    p = mpnn.unpack(theta); p["w_out"][:] = 0.3; theta = mpnn.pack(p)
    loss, g = mpnn.loss_and_grad(theta, gr, vt)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(theta), 60, replace=False)
    idx = np.unique(np.concatenate([idx, [len(theta) - 1, 0, 200, 300]]))
    for k in idx:
        h = 1e-6 * max(1.0, abs(theta[k]))
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        lp, _ = mpnn.loss_and_grad(tp, gr, vt)
        lm, _ = mpnn.loss_and_grad(tm, gr, vt)
        fd = (lp - lm) / (2 * h)
        # central differences: O(h^2) truncation plus relu kinks crossed within h
        assert abs(fd - g[k]) <= 1e-4 * abs(fd) + 1e-9, (k, fd, g[k])


def test_gate_gradients_are_exercised_and_planted_error_is_caught():
    gr, vt = small_graph(2)
    theta = mpnn.init_params(5)
    p = mpnn.unpack(theta); p["w_out"][:] = 0.3; theta = mpnn.pack(p)
    loss, g = mpnn.loss_and_grad(theta, gr, vt)
    names = [n for n, _ in mpnn.param_layout()]
    off = {}
    o = 0
    for n, s in mpnn.param_layout():
        off[n] = (o, o + int(np.prod(s))); o += int(np.prod(s))
    a0, a1 = off["ga1"]
    assert np.abs(g[a0:a1]).max() > 0
    # planted mistake: gradient of ga1 without the sigmoid derivative factor
    k = a0 + 3
    h = 1e-6
    tp = theta.copy(); tp[k] += h
    tm = theta.copy(); tm[k] -= h
    fd = (mpnn.loss_and_grad(tp, gr, vt)[0] - mpnn.loss_and_grad(tm, gr, vt)[0]) / (2 * h)
    gate_wrong = g[k] / 0.25      # drop the gate*(1-gate) <= 1/4 factor (rough)
    assert abs(fd - g[k]) < 1e-4 * abs(fd) + 1e-12
    assert abs(fd - gate_wrong) > 1e-3 * abs(fd)
    assert names[0] == "W_in"


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 3.14159265, -2.5e-3])
    r = mpnn.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 + 2 ** -7 or r[1] == 1.0
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
