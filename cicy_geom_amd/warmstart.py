"""Learned warm start for the DEA solve (binding of the device network in dea_mpnn.hip).

A 3-layer message-passing network over the conductance graph (node coordinates and
Jacobi-normalised log-conductance edges, 32 channels, BF16 weights, FP32 accumulation)
predicts a correction of the reference's linear ramp guess
(Initial_guess_for_CG, Direct_Electrifying.cpp:1396-1470).  It is trained on the device
(forward, backward and Adam are HIP kernels) on solves that the solver itself produces.
Zero-guess CG is the fallback and the yardstick: ``warm_start_gain`` solves every held-out
system both ways and reports the iterations saved.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import dea
from .network import DEASystem

H, F, LAYERS = 32, 4, 3


def param_layout():
    out = [("W_in", (H, F)), ("b_in", (H,))]
    for l in range(LAYERS):
        out += [(f"Ws{l}", (H, H)), (f"Wn{l}", (H, H)), (f"bias{l}", (H,)),
                (f"ga{l}", (H,)), (f"gc{l}", (H,))]
    out += [("w_out", (H,)), ("b_out", (1,))]
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """Scaled-normal weights (1/sqrt(fan_in)), small readout, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in param_layout():
        if name.startswith("W"):
            parts.append(rng.standard_normal(shape) * np.sqrt(1.0 / shape[1]))
        elif name.startswith("ga"):
            parts.append(rng.standard_normal(shape) * 0.5)
        elif name == "w_out":
            parts.append(rng.standard_normal(shape) * 0.01)
        else:
            parts.append(np.zeros(shape))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


class Graph:
    """Device graph of one assembled system (cicy_graph)."""

    def __init__(self, ctx: dea.Context, s: DEASystem, box: float, direction: int,
                 v_app: float, training: bool = False):
        if s.node_xyz is None:
            raise ValueError("assemble the system with with_xyz=True")
        self.ctx = ctx
        self.n = s.n
        h = ctypes.c_void_p()
        L = dea.lib()
        xyz = np.ascontiguousarray(s.node_xyz, np.float64).reshape(-1)
        ctx._check(L.cicy_graph_create(ctx._h, s.n, s.nnz, dea._i64(s.row_ptr),
                                       dea._i64(s.col_ind), dea._f64(s.values),
                                       dea._f64(s.diag), dea._f64(s.b), xyz, box, direction,
                                       v_app, 1 if training else 0, ctypes.byref(h)),
                   "cicy_graph_create")
        self._h = h
        ctx._adopt(self)

    def set_target(self, v_true):
        self.ctx._check(dea.lib().cicy_graph_set_target(self._h, dea._f64(v_true)),
                        "cicy_graph_set_target")

    def close(self):
        if getattr(self, "_h", None):
            dea.lib().cicy_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        if not dea._finalizing():
            self.close()


class WarmStartNet:
    """Parameters and Adam state on the device (cicy_mpnn)."""

    def __init__(self, ctx: dea.Context, theta: np.ndarray | None = None, seed: int = 0):
        self.ctx = ctx
        L = dea.lib()
        self.n_params = int(L.cicy_mpnn_num_params())
        th = init_params(seed) if theta is None else np.ascontiguousarray(theta, np.float32)
        assert th.size == self.n_params
        h = ctypes.c_void_p()
        ctx._check(L.cicy_mpnn_create(ctx._h, th.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.byref(h)), "cicy_mpnn_create")
        self._h = h
        ctx._adopt(self)

    def close(self):
        if getattr(self, "_h", None):
            dea.lib().cicy_mpnn_destroy(self._h)
            self._h = None

    def __del__(self):
        if not dea._finalizing():
            self.close()

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self.ctx._check(dea.lib().cicy_mpnn_get_params(self._h, out), "get_params")
        return out

    @params.setter
    def params(self, theta):
        self.ctx._check(dea.lib().cicy_mpnn_set_params(
            self._h, np.ascontiguousarray(theta, np.float32)), "set_params")

    def grads(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self.ctx._check(dea.lib().cicy_mpnn_get_grads(self._h, out), "get_grads")
        return out

    def zero_grads(self):
        self.ctx._check(dea.lib().cicy_mpnn_zero_grads(self._h), "zero_grads")

    def forward(self, g: Graph, want_loss: bool = False):
        v = np.zeros(g.n)
        loss = ctypes.c_double()
        self.ctx._check(dea.lib().cicy_mpnn_forward(
            self._h, g._h, v.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(loss) if want_loss else None), "cicy_mpnn_forward")
        return (v, loss.value) if want_loss else v

    def loss_and_grad(self, g: Graph):
        """Forward with loss then backward; gradients accumulate on the device (the
        readout part is accumulated by the forward)."""
        _, loss = self.forward(g, want_loss=True)
        self.ctx._check(dea.lib().cicy_mpnn_backward(self._h, g._h), "cicy_mpnn_backward")
        return loss

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, grad_scale=1.0):
        self.ctx._check(dea.lib().cicy_mpnn_adam_step(self._h, lr, beta1, beta2, eps,
                                                      grad_scale), "cicy_mpnn_adam_step")

    def warm_start(self, g: Graph, sysd: dea.System) -> float:
        """x0 := prediction and r0 := b - A x0 inside the device system; returns ms."""
        ms = ctypes.c_double()
        self.ctx._check(dea.lib().cicy_mpnn_warm_start(self._h, g._h, sysd._h,
                                                       ctypes.byref(ms)),
                        "cicy_mpnn_warm_start")
        return ms.value

    def train(self, graphs, epochs=60, lr=3e-3, lr_final=3e-4, log=None):
        """Adam over the training graphs (one step per graph), cosine-decayed rate."""
        hist = []
        steps = epochs * len(graphs)
        k = 0
        t0 = time.perf_counter()
        for ep in range(epochs):
            tot = 0.0
            for g in graphs:
                frac = k / max(1, steps - 1)
                rate = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
                tot += self.loss_and_grad(g)
                self.adam_step(rate)
                k += 1
            hist.append(tot / len(graphs))
            if log and (ep % 10 == 0 or ep == epochs - 1):
                log(f"[warmstart] epoch {ep} loss {hist[-1]:.3e} "
                    f"({time.perf_counter() - t0:.1f}s)")
        return hist


def training_set(ctx: dea.Context, nets_and_systems, tol=1e-12):
    """Solve each (system, box, direction) with the FP64 GPU solver and build training
    graphs with those voltages as targets."""
    graphs = []
    for s, box, direction, v_app in nets_and_systems:
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        sysd.set_rhs(s.b, None, s.b)
        st = sysd.solve(dea.METHOD_PCG_F64, tol, 0, dea.REL_B)
        x, _ = sysd.get()
        sysd.close()
        if st.status != dea.CICY_OK:
            continue
        g = Graph(ctx, s, box, direction, v_app, training=True)
        g.set_target(x)
        graphs.append(g)
    return graphs
