#!/usr/bin/env python3
"""Benchmark: DEA conductance solves/s to ||b - A x|| / ||b|| <= 1e-10 on one MI355X.

Workload: the 8 um RVE with 160k straight CNTs and 2k square nanoplatelets (seed
3_000_000), largest cluster percolating along x, R_flag=1 resistance solve (family := x),
zero initial guess.  The system's SSS arrays (the reference's row_ptr, col_ind, values,
diagonal, with int64 indices) and b are uploaded to HBM once, before anything is timed.  A
"step" is one new system, as the reference solves each (cluster, direction) system once:
create it from the arrays in HBM (checks and band layout on the device), set its right-hand
side, build the method's set-up on the device (condensation, series rounds, core layout,
plate-block inverses), solve it (x stays in HBM) and destroy it.  The rate from host arrays
(upload and host layout inside the step) and the rate of re-solving one prepared system are
reported beside it.

Methods (--methods): plain FP64 Jacobi-PCG on the GPU (the reference's algorithm, always
timed as the reference line), the condensed solves (exact series elimination down to a
core system, back-substitution) with the core solved by the two-kernel FP64 Jacobi-PCG
(condensed), the register-resident CG (condensed_resident) or the persistent
block-preconditioned CG with one dense block per nanoplatelet (condensed_block), the
register-resident CG on the full system, the mixed-precision FP32/FP64 solver, and the
learned warm start followed by a solver (network inference inside the step; needs trained
weights).  Methods that iterate on the full system (pcg_f64, learned_f64, mixed, resident)
time --plain-steps systems (about 13 s each here), the others --steps.  The mixed and
resident methods build their layouts from host arrays, so their steps start from those.
`value` is the fastest converged method's new systems per second; every method's true
residual ||b - A x|| / ||b|| is checked with the full system's FP64 SpMV and reported.
Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MAX_ITER = 400_000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


WEIGHTS = os.path.join(ROOT, "cicy_geom_amd", "weights", "warmstart.npz")


def time_solves(sysd, method, steps, warmup, tol, s=None, model=None, graph=None):
    """Wall time of `steps` solves on resident data (after `warmup` untimed ones).  With a
    model, every solve starts by writing the network's prediction into x0/r0."""
    from cicy_geom_amd import dea
    stats = []

    def one():
        # the zero-guess right-hand side was uploaded once; the learned method writes its
        # prediction into x0 and r0 = b - A x0 on the device as part of the solve
        if model is not None:
            model.warm_start(graph, sysd)
        return sysd.solve(method, tol, MAX_ITER, dea.REL_B)

    for _ in range(warmup):
        one()
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(one())
    wall = time.perf_counter() - t0
    return wall, stats


def time_fresh(ctx, s, method, steps, warmup, tol, plates=None, model=None, box=None,
               dev=None):
    """`steps` solves of the system as a new one each time (after `warmup` untimed ones).

    With `dev` (the reference's SSS arrays and b already in device memory, uploaded once
    before the timed steps): the system is created from them on the device
    (cicy_sys_create_device: checks and layout), its right-hand side set from device memory,
    the plate row groups given, the method's set-up built on the device (cicy_sys_prepare:
    eliminations, core layout, block inverses), the learned guess written when a model is
    given, the system solved (x stays in HBM) and destroyed.  Without `dev` the same from the
    host arrays (cicy_sys_create: upload and host layout, as a caller with host vectors
    sees it).  Every step is timed whole; returns per-step timings (s) and the stats."""
    from cicy_geom_amd import dea, warmstart
    out = []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        if dev is not None:
            sysd = ctx.system_device(s.n, len(s.values), dev["row_ptr"], dev["col_ind"],
                                     dev["values"], dev["diag"])
            sysd.set_rhs_device(dev["b"], None, dev["b"])
        else:
            sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
            sysd.set_rhs(s.b, None, s.b)
        if method == dea.METHOD_CONDENSED_BLOCK and plates is not None:
            sysd.set_blocks(plates)
        t1 = time.perf_counter()
        sysd.prepare(method)
        t2 = time.perf_counter()
        graph = None
        if model is not None:
            graph = warmstart.Graph(ctx, s, box, 0, 1.0)
            model.warm_start(graph, sysd)
        st = sysd.solve(method, tol, MAX_ITER, dea.REL_B)
        t3 = time.perf_counter()
        info = {k: sysd.size(k) for k in ("condensed_rows", "core_rows", "core_nnz",
                                          "series_rounds", "core_blocks", "core_block_rows",
                                          "core_groups", "core_block_entries", "format",
                                          "resident_rows")}
        if graph is not None:
            graph.close()
        sysd.close()
        t4 = time.perf_counter()
        if i >= warmup:
            out.append(dict(create_s=t1 - t0, prepare_s=t2 - t1, solve_s=t3 - t2,
                            destroy_s=t4 - t3, total_s=t4 - t0, stats=st, info=info))
    return out


def cpu_baseline(s, iters_needed, budget_s=12.0):
    """The reference's CPU algorithm (oracle restatement, 1 thread) on the same system:
    a bounded number of PCG iterations, timed, scaled to a full solve."""
    import oracle
    k = 5
    t = 0.0
    while True:
        t0 = time.perf_counter()
        oracle.pcg_sss(s.nodes, s.reserved, 0.0, s.row_ptr, s.col_ind, s.values, s.diag, s.b,
                       s.vef, np.zeros(s.n), max_iter=k)
        t = time.perf_counter() - t0
        if t > budget_s / 4 or k >= 4096:
            break
        k = int(k * max(2.0, min(8.0, budget_s / 4 / max(t, 1e-3))))
    t_iter = t / k
    solve_s = t_iter * iters_needed
    return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{k} Jacobi-PCG iterations of the same N={s.n} system in "
                      f"{t:.2f} s (oracle/dea_oracle.c, -O3 -ffp-contract=off, 1 thread, "
                      f"{cpu_model()}), scaled by the {iters_needed} iterations the FP64 GPU "
                      f"solve needed",
            "extrapolated": True, "iterations_timed": k, "iterations_scaled_to": iters_needed,
            "s_per_iter": t_iter, "host_nproc": os.cpu_count()}


# Contact search workload: the reference's defaults (input.dat) -- 0.01 um steps,
# 0.05 um background grid, van der Waals 0.00034 um, tunnelling 0.0018 um, min_points 5 --
# on tubes of radius 0.004-0.006 um at about 8 % volume.
CONTACT_CUTOFFS = (0.00034, 0.0018, 5)


def contact_args(box):
    return dict(window_min=[0.0, 0.0, 0.0], window_len=[box] * 3, gs_min=[0.05] * 3,
                overlap_cnt=2 * 0.006 + CONTACT_CUTOFFS[1], overlap_gnp=0.0)


def bench_contacts(args):
    """One step = one full contact search + labelling (cicy_cnt_contacts) of a synthetic
    wavy-tube network with its inputs in host memory, as the reference's call receives them;
    value = points per second over the K timed steps."""
    from cicy_geom_amd import dea, network
    box = args.contacts_box
    n_tubes = int(10_000 * (box / 2.0) ** 3)
    t0 = time.perf_counter()
    xyz, radii, structure = network.wavy_tubes(11, n_tubes, box, n_min=40, n_max=120)
    log(f"[bench] contacts: {len(xyz)} points in {len(structure)} tubes, {box} um cube, "
        f"built in {time.perf_counter() - t0:.1f}s")
    g = contact_args(box)
    ctx = dea.Context(0)
    off = np.concatenate([[0], np.cumsum([len(t) for t in structure])]).astype(np.int64)
    tubes = (off, np.arange(off[-1], dtype=np.int64))  # consecutive ids per tube

    def step():
        return ctx.cnt_contacts(xyz, radii, tubes, g["window_min"], g["window_len"],
                                g["gs_min"], g["overlap_cnt"], g["overlap_gnp"],
                                *CONTACT_CUTOFFS, as_arrays=True)
    for _ in range(args.warmup):
        res = step()
    times, stats = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        res = step()
        times.append(time.perf_counter() - t)
        stats.append(res["stats"])
        log(f"[bench] step {i}: {times[-1] * 1e3:.1f} ms {res['stats']}")
    ms = 1e3 * sum(times) / len(times)
    mean = {k: sum(s_[k] for s_ in stats) / len(stats) for k in stats[0]}
    cpu = None
    if not args.no_cpu:
        from oracle import hk
        sb = 2.0
        sx, sr, ss = network.wavy_tubes(11, int(10_000 * (sb / 2.0) ** 3), sb, n_min=40,
                                        n_max=120)
        gb = contact_args(sb)
        t = time.perf_counter()
        hk.cnt_clusters(gb["window_min"], gb["window_len"], gb["gs_min"], gb["overlap_cnt"],
                        gb["overlap_gnp"], CONTACT_CUTOFFS, sr, ss, sx)
        dt = time.perf_counter() - t
        cpu = {"value": len(sx) / dt, "unit": "points/s", "cores": 1, "kind": "port",
               "sample": f"oracle/hk.py (numpy per subregion, Python bookkeeping) on the same "
                         f"generator at a {sb} um cube ({len(sx)} points): {dt:.2f} s"}
    ctx.close()
    print(json.dumps({
        "metric": "contact_search_points_per_s", "value": len(xyz) / (ms * 1e-3),
        "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"tube contact search + HK labels, {len(xyz)} points, "
                               f"{len(structure)} tubes, {box} um cube, 0.05 um grid",
                   "points": len(xyz), "tubes": len(structure)},
        "stages_ms": {k: mean[k] for k in ("kernel_ms", "host_regions_ms", "device_ms",
                                            "host_labels_ms")},
        "counts": {k: mean[k] for k in ("subregions", "slots", "pair_tests", "contacts")},
        "labels": int(res["n_labels"]), "junctions": int(len(res["junctions"])),
        "cpu_baseline": cpu}), flush=True)


def bench_backbone(args):
    """Backbone extraction of the configuration's largest x-percolating cluster: the
    unit-resistor (R_flag = 0) system solved with the block method, then the device scan
    (zero current and tube runs, cicy_backbone_scan, inputs in host memory) timed over K
    calls, then the host pruning (cicy_net_backbone) with the scan's results, once (it edits
    the cluster)."""
    from cicy_geom_amd import dea, network
    t0 = time.perf_counter()
    net = network.generate(args.config)
    k = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    s0 = net.assemble(k, 0, net.cluster(k)[1])
    plates = network.plate_row_ranges(net, k, s0)
    log(f"[bench] backbone: config {args.config}, N={s0.n} built in "
        f"{time.perf_counter() - t0:.1f}s")
    ctx = dea.Context(0)
    sysd = ctx.system(s0.row_ptr, s0.col_ind, s0.values, s0.diag)
    sysd.set_rhs(s0.b, None, s0.b)
    sysd.set_blocks(plates)
    st = sysd.solve(dea.METHOD_CONDENSED_BLOCK, args.tol, 0, dea.REL_B)
    volt = np.concatenate([s0.vef, sysd.get()[0]])
    sysd.close()
    t = time.perf_counter()
    off, nodes, pa, pb = net.backbone_scan_inputs(k, s0)
    inputs_s = time.perf_counter() - t
    for _ in range(args.warmup):
        ctx.backbone_scan(volt, off, nodes, pa, pb)
    t = time.perf_counter()
    for _ in range(args.steps):
        zero, seg = ctx.backbone_scan(volt, off, nodes, pa, pb)
    scan_ms = 1e3 * (time.perf_counter() - t) / args.steps
    t = time.perf_counter()
    vol = net.backbone(k, s0, volt, zero=zero, segments=seg)
    prune_s = time.perf_counter() - t
    ctx.close()
    kept = int((seg[:, 0] >= 0).sum())
    print(json.dumps({
        "metric": "backbone_scan_ms", "value": scan_ms, "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"backbone of config {args.config}'s largest cluster, "
                               f"R_flag=0, N={s0.n}", "tubes": int(len(off) - 1),
                   "element_points": int(off[-1]), "pairs": int(len(pa))},
        "unit_solve": {"iters": int(st.iters), "ms": st.solve_ms,
                       "true_rel_res": st.true_rel_res},
        "scan_inputs_s": inputs_s, "host_pruning_s": prune_s, "zero_current": zero,
        "tubes_with_backbone": kept, "volumes": {k_: float(v) for k_, v in vol.items()}}),
        flush=True)


def bench_assembly(args):
    """Device assembly of the configuration's largest x-percolating cluster (R_flag = 1,
    family x, boundary lists of x) from the network held in HBM (cicy_dnet_assemble), timed
    over K calls; then the same system carried on into the solver without leaving the device
    (set-up + block-method solve), and the host assembler (cicy_net_assemble) timed once as
    the CPU baseline."""
    from cicy_geom_amd import dea, network
    t0 = time.perf_counter()
    net = network.generate(args.config)
    k = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    gen_s = time.perf_counter() - t0
    ctx = dea.Context(0)
    t = time.perf_counter()
    dn = net.device(ctx)
    copy_s = time.perf_counter() - t
    for _ in range(args.warmup):
        dn.assemble(k, 1, 0, None, 0).close()
    t = time.perf_counter()
    for _ in range(args.steps):
        a = dn.assemble(k, 1, 0, None, 0)
        a.close()
    asm_ms = 1e3 * (time.perf_counter() - t) / args.steps
    a = dn.assemble(k, 1, 0, None, 0)
    n, nnz, nblocks = a.n, a.nnz, len(a.blocks)
    a.close()
    # network in HBM -> solution in HBM: assembly, system set-up, block-method solve
    e2e = []
    for i in range(args.warmup + args.steps):
        t = time.perf_counter()
        a = dn.assemble(k, 1, 0, None, 0)
        sd = a.system()
        if len(a.blocks):
            sd.set_blocks(a.blocks)
        st = sd.solve(dea.METHOD_CONDENSED_BLOCK, args.tol, 0, dea.REL_B)
        sd.close()
        a.close()
        if i >= args.warmup:
            e2e.append(time.perf_counter() - t)
        if st.status != dea.CICY_OK:
            raise RuntimeError(f"solve after device assembly: {st.as_dict()}")
    e2e_ms = 1e3 * float(np.mean(e2e))
    t = time.perf_counter()
    net.assemble(k, 1, 0, direction=0)
    host_s = time.perf_counter() - t
    dn.close()
    ctx.close()
    print(json.dumps({
        "metric": "DEA system assemblies/s (device, network in HBM)",
        "value": 1e3 / asm_ms, "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": asm_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"assembly of config {args.config}'s largest x-percolating "
                               f"cluster, R_flag=1, N={n}, nnz={nnz}, {nblocks} plates"},
        "network_to_solution_ms": e2e_ms, "network_to_solution_per_s": 1e3 / e2e_ms,
        "network_copy_s": copy_s, "generate_s": gen_s,
        "cpu_baseline": {"value": 1.0 / host_s, "unit": "systems/s", "cores": 1,
                         "kind": "port", "sample": "one host assembly (cicy_net_assemble: "
                         "triplet list, stable sort, fold) of the same cluster"}}), flush=True)


def batch_realisation(k, keep_net=False):
    """Systems of realisation k of the batch workload (configuration 4; seed 4_000_000 + k):
    U[2k, 40k] tubes at the density of the 5 um / 40k-tube RVE, the largest cluster
    percolating along each direction, R_flag = 1 (family := direction).  With keep_net,
    also the network and the (cluster, direction) of each system."""
    from cicy_geom_amd import network
    out, jobs = [], []
    net = network.generate(4, realisation=k)
    for d in range(3):
        ks = net.percolating_clusters(d)
        if ks:
            c = max(ks, key=lambda q: net.cluster(q)[0])
            out.append(net.assemble(c, 1, d, direction=d))
            jobs.append((c, d))
    if keep_net:
        return out, net, jobs
    net.close()
    return out


def bench_batch(args):
    """The batch workload (configuration 4): every percolating direction of R realisations
    with U[2k, 40k] tubes, solved to ||b - A x|| <= 1e-10 ||b|| from the zero guess.
      (a) all R realisations with the block-preconditioned condensed solve, one new system
          after another, each timed from its host arrays (upload, layout, set-up, solve,
          read-back); generation is outside the timed region;
      (b) the first --batch-pcg-realisations of them in one batch (cicy_batch: the systems
          concatenated, per-system scalars and stopping tests in shared launches, FP64
          Jacobi-PCG, the reference's algorithm), timed from the host arrays likewise;
      (c) the same systems from each network held in HBM (copied once per realisation,
          not timed): device assembly, set-up and solve (network_to_solution_device).
    value = systems / s of (a) over all R realisations."""
    from cicy_geom_amd import dea
    ctx = dea.Context(0)
    t_gen, wall, count, iters, worst, rows = 0.0, 0.0, 0, 0, 0.0, []
    dwall, dcopy, dcount, dworst = 0.0, 0.0, 0, 0.0
    kept = []
    for k in range(args.batch_realisations):
        t = time.perf_counter()
        systems, net, jobs = batch_realisation(k, keep_net=True)
        t_gen += time.perf_counter() - t
        # (c) the same systems from the network in HBM: device assembly, set-up and solve
        t = time.perf_counter()
        dn = net.device(ctx)
        dcopy += time.perf_counter() - t
        for c, d in jobs:
            t = time.perf_counter()
            a = dn.assemble(c, 1, d, None, d)
            sysd = a.system()
            if len(a.blocks):
                sysd.set_blocks(a.blocks)
            st = sysd.solve(dea.METHOD_CONDENSED_BLOCK, args.tol, MAX_ITER, dea.REL_B)
            sysd.close()
            a.close()
            dwall += time.perf_counter() - t
            dcount += 1
            dworst = max(dworst, st.true_rel_res)
            if st.status != dea.CICY_OK:
                log(f"[bench] WARNING device-assembled system of realisation {k} did not "
                    "converge")
        dn.close()
        net.close()
        if k < args.batch_pcg_realisations:
            kept.extend(systems)
        for s in systems:
            t = time.perf_counter()
            sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
            sysd.set_rhs(s.b, None, s.b)
            st = sysd.solve(dea.METHOD_CONDENSED_BLOCK, args.tol, MAX_ITER, dea.REL_B)
            sysd.get()
            sysd.close()
            wall += time.perf_counter() - t
            count += 1
            iters += st.iters
            rows.append(s.n)
            worst = max(worst, st.true_rel_res)
            if st.status != dea.CICY_OK:
                log(f"[bench] WARNING condensed system N={s.n} did not converge")
        if (k + 1) % 16 == 0:
            log(f"[bench] batch: {k + 1} of {args.batch_realisations} realisations, {count} "
                f"systems, {wall:.1f} s of solves ({t_gen:.1f} s generating)")
    out = {"condensed_block_sequential": dict(
        systems=count, seconds=wall, solves_per_s=count / wall, core_iters=iters,
        worst_true_rel_res=worst, rows_min=min(rows), rows_max=max(rows),
        rows_total=int(sum(rows))),
        "network_to_solution_device": dict(
            systems=dcount, seconds=dwall, solves_per_s=dcount / dwall if dwall else None,
            network_copy_s=dcopy, worst_true_rel_res=dworst,
            step="device assembly + set-up + block-method solve, network in HBM")}
    log(f"[bench] batch network_to_solution_device: {out['network_to_solution_device']}")
    log(f"[bench] batch condensed_block: {out['condensed_block_sequential']}")
    if kept and not args.no_plain:
        t = time.perf_counter()
        B = ctx.batch([(s.row_ptr, s.col_ind, s.values, s.diag) for s in kept])
        for i, s in enumerate(kept):
            B.set_rhs(i, s.b, None, s.b)
        t1 = time.perf_counter()
        stats, dev_ms = B.solve(dea.METHOD_PCG_F64, args.tol, 0, dea.REL_B)
        for i in range(len(kept)):
            B.get(i)
        pw = time.perf_counter() - t
        create_s = t1 - t
        B.close()
        cw = 0.0
        for s in kept:   # the same systems with the condensed solve, for the comparison
            t2 = time.perf_counter()
            sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
            sysd.set_rhs(s.b, None, s.b)
            sysd.solve(dea.METHOD_CONDENSED_BLOCK, args.tol, MAX_ITER, dea.REL_B)
            sysd.get()
            sysd.close()
            cw += time.perf_counter() - t2
        out["batched_pcg_f64"] = dict(systems=len(kept), seconds=pw, create_s=create_s,
                                      solves_per_s=len(kept) / pw,
                                      iters=int(sum(x.iters for x in stats)),
                                      max_iters=int(max(x.iters for x in stats)),
                                      worst_true_rel_res=max(x.true_rel_res for x in stats),
                                      converged=int(sum(x.status == dea.CICY_OK
                                                        for x in stats)),
                                      condensed_block_same_systems_s=cw,
                                      speedup_condensed_vs_batched=pw / cw)
        log(f"[bench] batch pcg_f64: {out['batched_pcg_f64']}")
    ctx.close()
    a = out["condensed_block_sequential"]
    print(json.dumps({
        "metric": "DEA solves/s to 1e-10 residual, batch workload (configuration 4)",
        "value": a["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": a["systems"], "warmup": 0, "ms_per_step": 1e3 * a["seconds"] / a["systems"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"configuration 4: {args.batch_realisations} realisations x "
                               f"percolating directions = {a['systems']} systems, U[2k,40k] "
                               f"tubes, N {a['rows_min']}..{a['rows_max']}",
                   "method": "condensed_block",
                   "step": "one system from host arrays: upload + set-up + solve + read-back"},
        "generation_s": t_gen, "methods": out}), flush=True)


def bench_netio(args):
    """Network files of the contact-search workload (5.5 M points in a 4 um cube): write and
    read cnt_struct/cnt_coordinates as csv (6 digits) and dat; host only."""
    import tempfile
    from cicy_geom_amd import netio, network
    box = args.contacts_box
    xyz, radii, structure = network.wavy_tubes(11, int(10_000 * (box / 2.0) ** 3), box,
                                               n_min=40, n_max=120)
    npts = [len(t) for t in structure]
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for fmt in ("csv", "dat"):
            t = time.perf_counter()
            netio.write_network(d, fmt, [0, 0, 0], [box] * 3, npts, radii, xyz, prec_cnt=6)
            w = time.perf_counter() - t
            size = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d)
                       if f.startswith("cnt_") and f.endswith(fmt))
            t = time.perf_counter()
            r = netio.read_cnts(d, fmt)
            rd = time.perf_counter() - t
            assert len(r["xyz"]) == len(xyz)
            out[fmt] = {"write_s": w, "read_s": rd, "bytes": size,
                        "read_points_per_s": len(xyz) / rd, "read_MB_per_s": size / rd / 1e6}
    print(json.dumps({
        "metric": "network_file_read_points_per_s", "value": out["dat"]["read_points_per_s"],
        "unit": "points/s", "higher_is_better": True, "data": "synthetic",
        "config": {"workload": f"cnt_struct + cnt_coordinates, {len(xyz)} points, "
                               f"{len(structure)} tubes"}, "formats": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed solves of the plain FP64 reference line (about 19 s each at "
                         "configuration 3)")
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rows-per-block", type=int, default=0,
                    help="skip the autotuner and use this row-block size")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 reference line (profiling runs of one method)")
    ap.add_argument("--methods",
                    default="pcg_f64,condensed_block,learned_block,learned_f64,mixed,"
                            "learned_mixed",
                    help="comma list from pcg_f64, mixed, resident, condensed, "
                         "condensed_resident, condensed_block, learned_f64, learned_mixed, "
                         "learned_block (the learned ones need trained weights)")
    ap.add_argument("--copy-ref", action="store_true",
                    help="first time a 1 GiB device copy (the yardstick of hardware-counter "
                         "traffic runs)")
    ap.add_argument("--profile-spmv", action="store_true",
                    help="only run the SpMV kernels (both precisions) and a 1 GiB reference "
                         "copy, for hardware-counter runs")
    ap.add_argument("--workload", default="dea",
                    choices=["dea", "contacts", "backbone", "netio", "batch", "assembly"],
                    help="dea: the conductance solve (the default line); contacts: the tube "
                         "contact search and cluster labels on synthetic wavy tubes; backbone: "
                         "the backbone scan and pruning of the configuration's largest "
                         "cluster; netio: reading and writing the network files; assembly: "
                         "the device assembly of the configuration's largest cluster")
    ap.add_argument("--batch-realisations", type=int, default=256,
                    help="realisations of the batch workload (configuration 4)")
    ap.add_argument("--batch-pcg-realisations", type=int, default=8,
                    help="of those, how many are also solved as one FP64 PCG batch")
    ap.add_argument("--contacts-box", type=float, default=4.0,
                    help="cube side (um) of the contact-search workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); multi-GPU "
              "runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    from cicy_geom_amd import dea, network

    if args.workload == "contacts":
        bench_contacts(args)
        return
    if args.workload == "backbone":
        bench_backbone(args)
        return
    if args.workload == "netio":
        bench_netio(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "assembly":
        bench_assembly(args)
        return
    t0 = time.perf_counter()
    net = network.generate(args.config)
    box = net.params.box
    k_cl = max(net.percolating_clusters(0), key=lambda c: net.cluster(c)[0])
    s = net.assemble(k_cl, 1, 0, direction=0, with_xyz=True)
    plates = network.plate_row_ranges(net, k_cl, s)
    n_gnp = net.count("gnps")
    n_cnt = net.count("tubes")
    del net
    log(f"[bench] config {args.config}: N={s.n} L={s.nnz} ({s.gnp_nodes} plate nodes in "
        f"{len(plates)} plates) built in {time.perf_counter() - t0:.1f}s")

    from cicy_geom_amd import autotune
    ctx = dea.Context(0)
    if args.copy_ref:
        cms = ctx.copy_bench(1 << 30, 3)
        log(f"[bench] copy 1 GiB: {cms:.3f} ms -> {2 * (1 << 30) / cms / 1e6:.0f} GB/s")
    if args.profile_spmv:
        reps = 20
        copy_ms = ctx.copy_bench(1 << 30, reps)
        log(f"[bench] copy 1 GiB: {copy_ms:.3f} ms -> {2 * (1 << 30) / copy_ms / 1e6:.0f} GB/s")
        if args.rows_per_block:
            ctx.set_rows_per_block(args.rows_per_block)
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        sysd.set_rhs(s.b, None, s.b)
        for prec in (64, 32):
            ms = sysd.spmv_bench(prec, reps)
            sv = 8 if prec == 64 else 4
            algo = network.spmv_bytes(s.n, s.nnz, sv=sv, si=4)
            log(f"[bench] SpMV fp{prec}: {ms * 1e3:.1f} us/launch, algorithmic {algo / 1e6:.1f} MB "
                f"-> {algo / ms / 1e6:.0f} GB/s ({sysd.size('format')=})")
        core = sysd.core()
        if core is not None:
            cn, cl = core.size("n"), core.size("nnz")
            ms = core.spmv_bench(64, 200)
            algo = network.spmv_bytes(cn, cl, sv=8, si=4)
            log(f"[bench] core SpMV fp64 ({cn} rows, {cl} entries): {ms * 1e3:.1f} us/launch, "
                f"algorithmic {algo / 1e6:.2f} MB -> {algo / ms / 1e6:.0f} GB/s")
        print(json.dumps({"copy_ms": copy_ms, "reps": reps, "n": s.n, "nnz": s.nnz}))
        sysd.close()
        ctx.close()
        return
    catalog = {"pcg_f64": (dea.METHOD_PCG_F64, "pcg_f64", False),
               "mixed": (dea.METHOD_MIXED, "mixed", False),
               "resident": (dea.METHOD_RESIDENT, "pcg_f64", False),
               "condensed": (dea.METHOD_CONDENSED, "condensed", False),
               "condensed_resident": (dea.METHOD_CONDENSED_RESIDENT, "condensed", False),
               "condensed_block": (dea.METHOD_CONDENSED_BLOCK, "condensed", False),
               "learned_f64": (dea.METHOD_PCG_F64, "pcg_f64", True),
               "learned_mixed": (dea.METHOD_MIXED, "mixed", True),
               "learned_block": (dea.METHOD_CONDENSED_BLOCK, "condensed", True)}
    full_system = {"pcg_f64", "learned_f64", "mixed", "learned_mixed", "resident"}
    wanted = [m for m in args.methods.split(",") if m]
    bad = [m for m in wanted if m not in catalog]
    if bad:
        sys.exit(f"unknown method(s) {bad}; choose from {sorted(catalog)}")
    if "pcg_f64" not in wanted and not args.no_plain:
        wanted.insert(0, "pcg_f64")          # the plain FP64 solve is always the reference line
    model = graph = None
    if os.path.exists(WEIGHTS) and not args.no_learned:
        from cicy_geom_amd import warmstart
        theta = np.load(WEIGHTS, allow_pickle=False)["theta"]
        model = warmstart.WarmStartNet(ctx, theta)
    plan = [(m,) + catalog[m] for m in wanted if model is not None or not catalog[m][2]]

    tuned = {}
    for name, method in (("pcg_f64", dea.METHOD_PCG_F64), ("mixed", dea.METHOD_MIXED),
                         ("condensed", dea.METHOD_CONDENSED)):
        if not any(p[2] == name for p in plan):
            continue
        if args.rows_per_block:
            tuned[name] = {"format": dea.FORMAT_AUTO, "rows_per_block": args.rows_per_block}
        else:
            t0 = time.perf_counter()
            # the condensed methods keep the automatic layout choice: the full system and
            # its core differ in structure (band-dominated vs junction-dominated)
            layouts = (dea.FORMAT_AUTO,) if name == "condensed" else None
            tuned[name] = autotune.tune(ctx, s, method, iters=96, cache_path=None,
                                        layouts=layouts, rows=(256, 512, 1024), log=log)
            log(f"[bench] autotune {name}: {autotune.LAYOUTS.get(tuned[name]['format'])} "
                f"rb={tuned[name]['rows_per_block']} ({time.perf_counter() - t0:.1f}s)")

    # the reference's arrays of the system, resident in HBM before any timed step
    dev = {k: ctx.upload(np.ascontiguousarray(a, dt)) for k, a, dt in
           (("row_ptr", s.row_ptr, np.int64), ("col_ind", s.col_ind, np.int64),
            ("values", s.values, np.float64), ("diag", s.diag, np.float64),
            ("b", s.b, np.float64))}
    host_only = {"mixed", "learned_mixed", "resident"}   # layouts built from host arrays
    results, systems, by_kind = {}, {}, {}
    for name, method, tkey, learned in plan:
        autotune.apply(ctx, tuned[tkey])
        steps, warmup = ((args.plain_steps, 0) if name in full_system
                         else (args.steps, args.warmup))
        # the metric's step: a new system each time, set-up inside the timed step, from the
        # arrays in HBM
        fresh = time_fresh(ctx, s, method, steps, 1 if name not in full_system else 0,
                           args.tol, plates, model if learned else None, box,
                           dev=None if name in host_only else dev)
        st = fresh[-1]["stats"]
        wall = sum(f["total_s"] for f in fresh)
        mean = {k: sum(f[k] for f in fresh) / len(fresh)
                for k in ("create_s", "prepare_s", "solve_s", "destroy_s", "total_s")}
        info = fresh[-1]["info"]
        results[name] = dict(solves_per_s=steps / wall, ms_per_solve=1e3 * wall / steps,
                             steps=steps,
                             iters=st.iters, refinements=st.refinements,
                             true_rel_res=st.true_rel_res, status=dea.STATUS_NAMES[st.status],
                             device_ms=st.solve_ms,
                             layout=autotune.LAYOUTS.get(info["format"]),
                             condensed_rows=info["condensed_rows"],
                             core_rows=info["core_rows"], core_nnz=info["core_nnz"],
                             series_rounds=info["series_rounds"],
                             core_blocks=info["core_blocks"],
                             core_block_rows=info["core_block_rows"],
                             core_block_entries=info["core_block_entries"],
                             core_groups=info["core_groups"],
                             step_breakdown_s=mean,
                             inputs="host arrays" if name in host_only else "device arrays",
                             setup_s=mean["create_s"] + mean["prepare_s"],
                             kernel_ms=sum(f["stats"].kernel_ms for f in fresh),
                             resident_rows_per_thread=info["resident_rows"],
                             rows_per_block=tuned[tkey]["rows_per_block"])
        if name not in full_system:
            # a new system from the host arrays each step (upload and host-side layout
            # included: the rate of a caller that hands over host vectors)
            hfresh = time_fresh(ctx, s, method, min(steps, 5), 1, args.tol, plates,
                                model if learned else None, box)
            results[name]["host_arrays_solves_per_s"] = len(hfresh) / sum(
                f["total_s"] for f in hfresh)
            results[name]["host_arrays_create_s"] = sum(
                f["create_s"] for f in hfresh) / len(hfresh)
            # the same method re-solving an already prepared, resident system (the rate a
            # caller gets when one system is solved many times)
            sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
            sysd.set_rhs(s.b, None, s.b)
            if method == dea.METHOD_CONDENSED_BLOCK:
                sysd.set_blocks(plates)
            sysd.prepare(method)
            graph = warmstart.Graph(ctx, s, box, 0, 1.0) if learned else None
            rwall, rstats = time_solves(sysd, method, steps, args.warmup, args.tol, s,
                                        model if learned else None, graph)
            results[name]["resident_solves_per_s"] = steps / rwall
            results[name]["resident_ms_per_solve"] = 1e3 * rwall / steps
            results[name]["resident_kernel_ms"] = sum(x.kernel_ms for x in rstats)
            if graph is not None:
                graph.close()
            systems[name] = sysd
            by_kind.setdefault(tkey, sysd)
        log(f"[bench] {name}: {results[name]}")
        if st.status != dea.CICY_OK:
            log(f"[bench] WARNING {name} did not converge")
    if "pcg_f64" not in by_kind and "pcg_f64" in results:
        sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
        sysd.set_rhs(s.b, None, s.b)
        systems["pcg_f64"] = sysd
        by_kind["pcg_f64"] = sysd

    # SpMV roofline (HIP events on the library's stream, back-to-back launches), next to a
    # plain 1 GiB device copy (read + write) as the achievable streaming rate
    n, L = s.n, s.nnz
    copy_gbs = 2 * (1 << 30) / (ctx.copy_bench(1 << 30, 5) * 1e-3) / 1e9
    read_gbs = (1 << 30) / (ctx.read_bench(1 << 30, 10) * 1e-3) / 1e9
    log(f"[bench] copy 1 GiB: {copy_gbs:.0f} GB/s; read 1 GiB: {read_gbs:.0f} GB/s")
    spmv = {}
    fp64_sys = next((v for k, v in by_kind.items() if k != "mixed"), None)
    for prec, sv, sysx in ((64, 8, fp64_sys), (32, 4, by_kind.get("mixed"))):
        if sysx is None:
            continue
        ms = sysx.spmv_bench(prec, 50)
        cold = sysx.spmv_bench_cold(prec, 20, 1 << 30)
        algo = network.spmv_bytes(n, L, sv=sv, si=4)
        spmv[prec] = dict(ms=ms, bytes=algo, gbs=algo / (ms * 1e-3) / 1e9, cold_ms=cold,
                          cold_gbs=algo / (cold * 1e-3) / 1e9)
        log(f"[bench] SpMV fp{prec}: {ms * 1e3:.1f} us back to back, {cold * 1e3:.1f} us with the "
            f"cache flushed; {algo / 1e6:.1f} MB algorithmic: {spmv[prec]['gbs']:.0f} / "
            f"{spmv[prec]['cold_gbs']:.0f} GB/s")

    plain = results.get("pcg_f64")
    head = "pcg_f64" if plain else next(iter(results))
    for cand in ("learned_f64", "mixed", "learned_mixed", "resident", "condensed",
                 "condensed_resident", "condensed_block", "learned_block"):
        if cand in results and cand != head and results[cand]["status"] == "ok" and \
                results[cand]["solves_per_s"] > results[head]["solves_per_s"]:
            head = cand
    value = results[head]["solves_per_s"]
    hp = 32 if catalog[head][1] == "mixed" else 64
    # whole-iteration rate of the plain FP64 solve (device time / iterations)
    it_bytes = network.pcg_iter_bytes(n, L)
    it_us = 1e3 * plain["device_ms"] / max(1, plain["iters"]) if plain else 0.0
    # roofline of the headline method's dominant kernel: the SSS product of the system its
    # CG iterates on (the full system, or the core of the condensed methods)
    roof = {"bound": "hbm", "achieved": spmv[hp]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": spmv[hp]["gbs"] / HBM_PEAK_GBS, "traffic": None,
            "kernel": "SSS SpMV (k_apply/k_apply_band) " +
                      ("fp64" if hp == 64 else "fp32 Jacobi-scaled") + " of the full system",
            "algorithmic_bytes": spmv[hp]["bytes"], "us_per_launch": spmv[hp]["ms"] * 1e3}
    kind = catalog[head][0]
    if kind == dea.METHOD_CONDENSED_BLOCK:
        # dominant kernel: the persistent block-preconditioned core CG (k_core_cg), one
        # cooperative launch per solve; its duration is the HIP-event time of its launches
        # inside the timed solves.  Algorithmic bytes per core iteration, counted as for a
        # matrix stored with 4-byte indices: the off-diagonal entries of both triangles
        # (column + value, 12 B), the diagonal with its row index (12 B per row), the FP32
        # dense block remainders (4 B per entry) and the vectors (x, r, u, w, p, s: 13
        # reads/writes of 8 B per row).  The traffic per iteration comes from the committed
        # counter runs of the same workload, scaled to this launch's iterations.
        r_ = results[head]
        cn, cl = r_["core_rows"], r_["core_nnz"]
        nb2 = r_["core_block_entries"]
        cib = 12 * 2 * cl + 12 * cn + 4 * nb2 + 8 * 13 * cn
        ksec = r_["kernel_ms"] * 1e-3 / max(1, r_["steps"])
        ach = cib * r_["iters"] / ksec / 1e9 if ksec > 0 else 0.0
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None,
                "kernel": f"k_core_cg: block-preconditioned core CG, {cn} rows, "
                          f"{r_['core_blocks']} plate blocks",
                "algorithmic_bytes": cib * r_["iters"], "algorithmic_bytes_per_iter": cib,
                "us_per_launch": ksec * 1e6, "iters_per_launch": r_["iters"],
                "us_per_iter": ksec * 1e6 / max(1, r_["iters"])}
        tpath = os.path.join(ROOT, "profiles", "core_block_6p0M_traffic.json")
        if args.config == 3 and os.path.exists(tpath):
            with open(tpath) as f:
                t = json.load(f)
            roof["traffic"] = t["bytes_per_iter"] * r_["iters"]
            roof["traffic_per_iter"] = t["bytes_per_iter"]
            roof["traffic_source"] = t["source"]
        log(f"[bench] k_core_cg: {roof['us_per_iter']:.1f} us per core iteration, "
            f"{cib / 1e6:.1f} MB algorithmic -> {ach:.0f} GB/s ({100 * roof['frac']:.1f} % of peak)")
    elif head.startswith("condensed") or kind == dea.METHOD_CONDENSED:
        core = systems[head].core()  # the resident system of the same method
        # dominant kernels: the core system's PCG iteration (k_apply<PCG> + k_update_f64);
        # its rate is the HIP-event device time of the timed solves over their iterations
        cn, cl = core.size("n"), core.size("nnz")
        cit = results[head]["device_ms"] * 1e3 / max(1, results[head]["iters"])
        cib = network.pcg_iter_bytes(cn, cl)
        roof = {"bound": "hbm", "achieved": cib / (cit * 1e-6) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": cib / (cit * 1e-6) / 1e9 / HBM_PEAK_GBS,
                "traffic": None,
                "kernel": f"PCG iteration (k_apply<PCG> + k_update_f64) of the {cn}-row core",
                "algorithmic_bytes": cib, "us_per_launch": cit}
    out = {
        "metric": "CG solves/s to 1e-10 residual at 4M nodes; learned+mixed vs plain FP64 GPU CG",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": results[head]["steps"],
        "warmup": 0 if head == "pcg_f64" else args.warmup,
        "ms_per_step": results[head]["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if hp == 64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"config {args.config}: {s.n}-node DEA system ({box:g} um RVE, "
                               f"{n_cnt} straight CNTs + {n_gnp} square plates, largest "
                               "x-percolating cluster, R_flag=1)",
                   "N": s.n, "L": s.nnz, "tol": args.tol, "residual": "||b-Ax||/||b||",
                   "method": head},
        "methods": results,
        "speedup_vs_plain_f64": value / plain["solves_per_s"] if plain else None,
        # each step is a new system from the reference's arrays in HBM: checks and layout,
        # set-up (eliminations, core layout, block inverses), solve (x in HBM), destroy
        "step": "new system from its SSS arrays resident in HBM: device layout + set-up + "
                "solve + destroy",
        "step_breakdown_s": results[head]["step_breakdown_s"],
        "setup_s": results[head]["setup_s"],
        "resident_solves_per_s": results[head].get("resident_solves_per_s"),
        "roofline": roof,
        "roofline_spmv_full_f64": {"achieved": spmv[64]["gbs"],
                                   "frac": spmv[64]["gbs"] / HBM_PEAK_GBS,
                                   "us": spmv[64]["ms"] * 1e3,
                                   "cold_achieved": spmv[64]["cold_gbs"],
                                   "cold_frac": spmv[64]["cold_gbs"] / HBM_PEAK_GBS,
                                   "cold_us": spmv[64]["cold_ms"] * 1e3,
                                   "algorithmic_bytes": spmv[64]["bytes"],
                                   "copy_1GiB_gbs": copy_gbs, "read_1GiB_gbs": read_gbs},
        "pcg_iteration_full": {"algorithmic_bytes": it_bytes, "us": it_us,
                               "gbs": it_bytes / (it_us * 1e-6) / 1e9 if it_us > 0 else None},
    }
    if not args.no_cpu and plain:
        out["cpu_baseline"] = cpu_baseline(s, plain["iters"])
        log(f"[bench] cpu baseline: {out['cpu_baseline']}")
    for sysd in systems.values():
        sysd.close()
    for a in dev.values():
        a.close()
    if model is not None:
        learned = {"weights": os.path.relpath(WEIGHTS, ROOT)}
        for lk, zk in (("learned_f64", "pcg_f64"), ("learned_mixed", "mixed"),
                       ("learned_block", "condensed_block")):
            if lk in results and zk in results:
                learned[f"iters_saved_{lk}_vs_{zk}"] = 1.0 - results[lk]["iters"] / results[zk]["iters"]
        out["learned"] = learned
        model.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
