"""Train the learned warm start on solves the solver itself generates, then assess it.

    python -m cicy_geom_amd.train_warmstart --n-train 64 --epochs 40 --out weights.npz

Training data: realisations of the batch workload's generator (seeds 4_000_000 + k, tube
counts drawn from [--min-cnt, --max-cnt] at the density of the 5 um / 40k-tube RVE), every
percolating direction of the largest cluster, R_flag = 1.  Targets are FP64 GPU solves to
||b - Ax|| <= 1e-12 ||b|| (the condensed solve; the plain FP64 PCG where a system has nothing
to condense).  Training runs on the device (forward, backward and Adam kernels).

Judging: held-out realisations (seeds 4_500_000 + k) are solved to ||b - Ax|| <= 1e-10 ||b||
with the plain FP64 PCG and with the condensed solve, each from the zero guess (the
fallback and yardstick), from the reference's ramp and from the learned guess; the iterations
of each are reported.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import dea, network, warmstart


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def systems_for(seed_base: int, count: int, min_cnt: int, max_cnt: int, rng_seed: int):
    rng = np.random.default_rng(rng_seed)
    out = []
    for k in range(count):
        ncnt = int(rng.integers(min_cnt, max_cnt + 1))
        p = network.config_params(4, 0, ncnt)
        p.seed = seed_base + k
        net = network.Network(p)
        for d in range(3):
            ks = net.percolating_clusters(d)
            if not ks:
                continue
            c = max(ks, key=lambda q: net.cluster(q)[0])
            s = net.assemble(c, 1, d, direction=d, with_xyz=True)
            out.append((s, p.box, d, 1.0))
        net.close()
        if (k + 1) % 50 == 0:
            log(f"[train] {k + 1} of {count} realisations generated ({len(out)} systems)")
    return out


def solve_targets(ctx, items, tol=1e-12):
    """FP64 solves to ||b - Ax|| <= tol ||b||: the voltages of each system (None if not
    converged) and the device milliseconds spent."""
    xs, ms = [], 0.0
    for i, (s, *_) in enumerate(items):
        if (i + 1) % 200 == 0:
            log(f"[train] {i + 1} of {len(items)} targets solved")
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        sysd.set_rhs(s.b, None, s.b)
        try:
            st = sysd.solve(dea.METHOD_CONDENSED, tol, 0, dea.REL_B)
        except dea.DEAError:
            st = sysd.solve(dea.METHOD_PCG_F64, tol, 0, dea.REL_B)
        x, _ = sysd.get()
        xs.append(x if st.status == dea.CICY_OK else None)
        ms += st.solve_ms
        sysd.close()
    return xs, ms


def assess(ctx, model, items, methods=(dea.METHOD_PCG_F64, dea.METHOD_CONDENSED), tol=1e-10):
    """Iterations (and solve ms) of each method from the zero guess, the reference's ramp
    and the learned guess, to ||b - A x|| <= tol ||b||."""
    rows = []
    for s, box, d, vapp in items:
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        g = warmstart.Graph(ctx, s, box, d, vapp)
        res = {"n": s.n}
        for method in methods:
            name = dea.METHOD_NAMES[method]
            sysd.set_rhs(s.b, None, s.b)
            st = sysd.solve(method, tol, 0, dea.REL_B)
            res[f"{name}/zero"] = (st.iters, st.solve_ms, st.status)
            sysd.set_rhs(s.r0 if s.has_guess else s.b, s.x0, s.b)
            st = sysd.solve(method, tol, 0, dea.REL_B)
            res[f"{name}/ramp"] = (st.iters, st.solve_ms, st.status)
            ms = model.warm_start(g, sysd)
            st = sysd.solve(method, tol, 0, dea.REL_B)
            res[f"{name}/learned"] = (st.iters, st.solve_ms + ms, st.status)
        g.close()
        sysd.close()
        rows.append(res)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=48)
    ap.add_argument("--n-test", type=int, default=6)
    ap.add_argument("--min-cnt", type=int, default=2000)
    ap.add_argument("--max-cnt", type=int, default=8000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="warmstart.npz")
    args = ap.parse_args(argv)

    ctx = dea.Context(0)
    t0 = time.perf_counter()
    train = systems_for(4_000_000, args.n_train, args.min_cnt, args.max_cnt, 1)
    log(f"[train] {len(train)} training systems generated in {time.perf_counter() - t0:.1f}s "
        f"(N {min(s.n for s, *_ in train)}..{max(s.n for s, *_ in train)})")
    xs, ms = solve_targets(ctx, train)
    log(f"[train] targets solved (FP64, 1e-12) in {ms / 1e3:.2f}s device time, "
        f"{sum(x is not None for x in xs)} of {len(xs)} converged")
    graphs = []
    for (s, box, d, vapp), x in zip(train, xs):
        if x is None:
            continue
        g = warmstart.Graph(ctx, s, box, d, vapp, training=True)
        g.set_target(x)
        graphs.append(g)
    model = warmstart.WarmStartNet(ctx, seed=args.seed)
    t0 = time.perf_counter()
    hist = model.train(graphs, epochs=args.epochs, lr=args.lr, log=log)
    train_s = time.perf_counter() - t0
    for g in graphs:
        g.close()
    test = systems_for(4_500_000, args.n_test, args.min_cnt, args.max_cnt, 2)
    rows = assess(ctx, model, test)
    summary = {"train_systems": len(graphs), "epochs": args.epochs, "train_seconds": train_s,
               "loss_first": hist[0], "loss_last": hist[-1]}
    for key in rows[0]:
        if key == "n":
            continue
        summary[f"iters {key}"] = int(sum(r[key][0] for r in rows))
    for r in rows:
        log("[assess] N=%d " % r["n"] + ", ".join(f"{k} {v[0]} it" for k, v in r.items()
                                                 if k != "n"))
    for name in ("pcg_f64", "condensed"):
        z, l = summary.get(f"iters {name}/zero"), summary.get(f"iters {name}/learned")
        if z:
            summary[f"iters_saved_{name}"] = 1.0 - l / z
    log("[assess] " + json.dumps(summary))
    np.savez(args.out, theta=model.params, loss=np.array(hist),
             summary=json.dumps(summary), assessment=json.dumps(rows))
    print(json.dumps(summary))
    model.close()
    ctx.close()


if __name__ == "__main__":
    main()
