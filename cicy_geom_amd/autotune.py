"""Autotuner for the solver's device layout and row-block size.

For a given system (or the largest system of a batch) every candidate (layout, rows per
work-group) is uploaded and timed on a fixed number of iterations of the chosen method
(tol = 0, so the solve runs exactly that many iterations; the device time comes from HIP
events inside the library).  The fastest candidate per (device, method, size bucket) is kept
in a JSON cache so later runs reuse it.  The candidates are the template instantiations the
library ships: layouts {blocked SSS, band + remainder} x rows per block {128, 256, 512, 1024}.
"""
from __future__ import annotations

import json
import math
import os

from . import dea

LAYOUTS = {dea.FORMAT_AUTO: "auto", dea.FORMAT_SSS: "blocked-sss",
           dea.FORMAT_BAND: "band+remainder"}
ROWS_PER_BLOCK = (128, 256, 512, 1024)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "cicy_geom_amd",
                             "autotune.json")


def _bucket(n: int) -> str:
    return f"2^{int(math.log2(max(n, 1)))}"


def time_candidate(ctx: dea.Context, s, fmt: int, rb: int, method: int, iters: int) -> float:
    """ms per iteration of `method` on system `s` with layout `fmt` and `rb` rows/block."""
    ctx.set_format(fmt)
    ctx.set_rows_per_block(rb)
    sysd = ctx.system(s.row_ptr, s.col_ind, s.values, s.diag)
    try:
        sysd.set_rhs(s.b, None, s.b)
        sysd.solve(method, 0.0, max(32, iters // 4), dea.REL_B)       # warm-up (graph capture)
        st = sysd.solve(method, 0.0, iters, dea.REL_B)
        return st.solve_ms / max(1, st.iters)
    finally:
        sysd.close()
        ctx.set_format(dea.FORMAT_AUTO)
        ctx.set_rows_per_block(0)


def tune(ctx: dea.Context, s, method: int = dea.METHOD_MIXED, iters: int = 128,
         cache_path: str | None = DEFAULT_CACHE, device_key: str = "gfx950",
         layouts=None, rows=ROWS_PER_BLOCK, log=None) -> dict:
    """Measure every candidate on `s` and return {"format", "rows_per_block", "table"};
    the winner is stored in the cache under (device, method, size bucket)."""
    layouts = layouts or (dea.FORMAT_SSS, dea.FORMAT_BAND)
    table = []
    for fmt in layouts:
        for rb in rows:
            ms = time_candidate(ctx, s, fmt, rb, method, iters)
            table.append({"format": fmt, "layout": LAYOUTS[fmt], "rows_per_block": rb,
                          "ms_per_iter": ms})
            if log:
                log(f"[autotune] {LAYOUTS[fmt]:>15} rb={rb:<5} {ms * 1e3:8.1f} us/iter")
    best = min(table, key=lambda t: t["ms_per_iter"])
    out = {"format": best["format"], "rows_per_block": best["rows_per_block"], "table": table}
    if cache_path:
        key = f"{device_key}/{method}/{_bucket(s.n)}"
        data = load_cache(cache_path)
        data[key] = {"format": best["format"], "rows_per_block": best["rows_per_block"],
                     "ms_per_iter": best["ms_per_iter"]}
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return out


def load_cache(cache_path: str = DEFAULT_CACHE) -> dict:
    try:
        with open(cache_path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def cached(n: int, method: int, cache_path: str = DEFAULT_CACHE, device_key: str = "gfx950"):
    """(format, rows_per_block) from the cache for this size bucket, or None."""
    hit = load_cache(cache_path).get(f"{device_key}/{method}/{_bucket(n)}")
    return (hit["format"], hit["rows_per_block"]) if hit else None


def apply(ctx: dea.Context, choice: dict | tuple):
    fmt, rb = (choice["format"], choice["rows_per_block"]) if isinstance(choice, dict) else choice
    ctx.set_format(fmt)
    ctx.set_rows_per_block(rb)
