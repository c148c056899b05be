"""Float64 restatement of the warm-start message-passing network -- TEST INFRASTRUCTURE ONLY.

Not a reference algorithm (the reference has no learned component; its warm start is the
linear ramp of Direct_Electrifying.cpp:1396-1470).  This module is the checker of the HIP
forward/backward kernels (cicy_geom_amd/csrc/dea_mpnn.hip): same graph, features,
parameter layout and loss, computed in float64 with plain numpy.  Its own gradient is
pinned by central finite differences (tests/test_mpnn_oracle.py).

Network (hidden H = 32, F = 4 node features, 3 message-passing layers):
  z0_i = W_in f_i + b_in                    h_i = relu(z0_i)
  per layer l:
    gate_ij = sigmoid(ga_l * e_ij + gc_l)   (per channel; e_ij = normalised log10 g_ij)
    agg_i   = sum_j w_ij * gate_ij * h_j    (w_ij = g_ij / d_i, Jacobi weights)
    z_i     = Ws_l h_i + Wn_l agg_i + bias_l
    h_i     = h_i + relu(z_i)
  y_i   = w_out . h_i + b_out
  V^_i  = V_app * (1 - u_i + y_i)           (a learned correction of the ramp)
  loss  = mean_i ((V^_i - V*_i) / V_app)^2
Features f_i = (u, v, w, beta): u the coordinate along the current direction over the
RVE side (the ramp coordinate), v, w the transverse ones, beta_i = b_i / (d_i V_app) the
pull of the high-voltage electrode.
"""
from __future__ import annotations

import numpy as np

H = 32
F = 4
LAYERS = 3
E_MU, E_SIGMA = -5.0, 2.5   # log10-conductance normalisation (g spans ~1e-10 .. 3e-3 S)


def param_layout():
    """(name, shape) in the order of the flat parameter vector."""
    out = [("W_in", (H, F)), ("b_in", (H,))]
    for l in range(LAYERS):
        out += [(f"Ws{l}", (H, H)), (f"Wn{l}", (H, H)), (f"bias{l}", (H,)),
                (f"ga{l}", (H,)), (f"gc{l}", (H,))]
    out += [("w_out", (H,)), ("b_out", (1,))]
    return out


def n_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unpack(theta):
    p, o = {}, 0
    for name, shape in param_layout():
        k = int(np.prod(shape))
        p[name] = np.asarray(theta[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return p


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1)
                           for name, _ in param_layout()])


def bf16_round(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.view(np.float32).astype(np.float64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_layout():
        if name.startswith("W"):
            fan_in = shape[1]
            p[name] = rng.standard_normal(shape) * np.sqrt(1.0 / fan_in)
        elif name.startswith("ga"):
            p[name] = rng.standard_normal(shape) * 0.5
        elif name == "w_out":
            p[name] = rng.standard_normal(shape) * 0.01
        else:
            p[name] = np.zeros(shape)
    return pack(p)


def graph_from_sss(row_ptr, col_ind, values, diag, b, node_xyz, box, direction, v_app):
    """Symmetric adjacency (both directions of every stored lower entry) with Jacobi
    weights and normalised log-conductance, plus node features and the ramp coordinate."""
    n = len(diag)
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    g = -np.asarray(values, np.float64)
    src = np.concatenate([rows, col_ind])          # receiving node i
    dst = np.concatenate([col_ind, rows])          # neighbour j
    gg = np.concatenate([g, g])
    order = np.lexsort((dst, src))
    src, dst, gg = src[order], dst[order], gg[order]
    nbr_ptr = np.zeros(n + 1, np.int64)
    np.add.at(nbr_ptr, src + 1, 1)
    nbr_ptr = np.cumsum(nbr_ptr)
    w = gg / np.asarray(diag)[src]
    e = (np.log10(gg) - E_MU) / E_SIGMA
    perm = [direction] + [a for a in range(3) if a != direction]
    xyz = np.asarray(node_xyz, np.float64)[:, perm] / box
    beta = np.asarray(b, np.float64) / (np.asarray(diag, np.float64) * v_app)
    feat = np.column_stack([xyz, beta])
    return dict(n=n, nbr_ptr=nbr_ptr, nbr=dst, w=w, e=e, feat=feat, u=xyz[:, 0], v_app=v_app,
                src=src)


def forward(theta, gr, keep=False):
    p = unpack(theta)
    src, dst, w, e = gr["src"], gr["nbr"], gr["w"], gr["e"]
    z0 = gr["feat"] @ p["W_in"].T + p["b_in"]
    h = np.maximum(z0, 0.0)
    cache = dict(z0=z0, hs=[h], aggs=[], zs=[], gates=[])
    for l in range(LAYERS):
        gate = 1.0 / (1.0 + np.exp(-(np.outer(e, p[f"ga{l}"]) + p[f"gc{l}"])))
        msg = (w[:, None] * gate) * h[dst]
        agg = np.zeros_like(h)
        np.add.at(agg, src, msg)
        z = h @ p[f"Ws{l}"].T + agg @ p[f"Wn{l}"].T + p[f"bias{l}"]
        h = h + np.maximum(z, 0.0)
        cache["aggs"].append(agg)
        cache["zs"].append(z)
        cache["gates"].append(gate)
        cache["hs"].append(h)
    y = h @ p["w_out"] + p["b_out"][0]
    vhat = gr["v_app"] * (1.0 - gr["u"] + y)
    return (vhat, cache) if keep else vhat


def loss_and_grad(theta, gr, v_true):
    """MSE of V^/V_app against V*/V_app and its gradient w.r.t. the flat parameters."""
    p = unpack(theta)
    vhat, c = forward(theta, gr, keep=True)
    n = gr["n"]
    d = (vhat - v_true) / gr["v_app"]
    loss = float(np.mean(d * d))
    g = {k: np.zeros_like(v) for k, v in p.items()}
    dy = 2.0 * d / n                      # dL/dy (V^/V_app = 1 - u + y)
    hL = c["hs"][-1]
    g["w_out"] = hL.T @ dy
    g["b_out"] = np.array([dy.sum()])
    dh = np.outer(dy, p["w_out"])
    src, dst, w, e = gr["src"], gr["nbr"], gr["w"], gr["e"]
    for l in reversed(range(LAYERS)):
        hin = c["hs"][l]
        z, agg, gate = c["zs"][l], c["aggs"][l], c["gates"][l]
        dz = dh * (z > 0)
        g[f"Ws{l}"] = dz.T @ hin
        g[f"Wn{l}"] = dz.T @ agg
        g[f"bias{l}"] = dz.sum(0)
        dagg = dz @ p[f"Wn{l}"]
        dhin = dh + dz @ p[f"Ws{l}"]
        # agg_i = sum_j w_ij gate_ij h_j
        dmsg = dagg[src]                                   # per edge (i <- j)
        np.add.at(dhin, dst, dmsg * w[:, None] * gate)
        dgate_pre = dmsg * w[:, None] * hin[dst] * gate * (1.0 - gate)
        g[f"ga{l}"] = (dgate_pre * e[:, None]).sum(0)
        g[f"gc{l}"] = dgate_pre.sum(0)
        dh = dhin
    dz0 = dh * (c["z0"] > 0)
    g["W_in"] = dz0.T @ gr["feat"]
    g["b_in"] = dz0.sum(0)
    return loss, pack(g)
