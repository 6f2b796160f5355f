"""Autotuner for the tile configurations of the FP64 GEMMs.

    python -m thcpp_amd.autotune [--out thcpp_amd/tuning/gemm_gfx950.json]

The GEMMs of the ISDF stages choose a tile configuration by a built-in rule
on their tile count (gemm.hip).  This tool measures, on the GPU it runs on,
every configuration (thc_gemm_time) on the GEMM shapes of three workloads --
the 32-cell uneven batch (24^3-48^3 grids, N_mu 128-1280) and the 64-atom
system at N_mu 1024 and 2048 -- for the two GEMMs whose shapes vary with the
system: ZC^T (N x N_mu x M, rows of Phi gathered, square epilogue) and the
FP64 solve (N_mu x N x N_mu).  Shapes are grouped by (ceil log2 m, n, k) and
kind; each group gets the configuration with the least summed time.

A GEMM alone is not the whole story (the ZC^T GEMM of a large grid keeps one
workgroup per CU so that the Cholesky of CC^T on the second stream finds free
CUs), so each candidate entry is then validated end to end, largest measured
gain first: the workloads with shapes in its bucket are timed (mixed and FP64
factorisations, or the batched call) with the entries kept so far, with and
without it, and it is kept only if they got faster.
The table, the per-configuration times and the validation timings are written
as JSON; thcpp_amd.api applies the committed table to every process.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import time

import torch

from . import _lib as L
from . import synthetic
from .api import TUNING_FILE, Context

CFG_NAMES = {1: "tiny", 2: "small", 3: "large", 4: "dtl", 5: "deepk"}
KIND_FLAGS = {0: 0, 1: 1 | 4}  # solve: plain; ZC^T: gathered operand + square epilogue
CONFIG_WORKLOADS = ("diamond64_n64_k1024", "diamond64_n64_k2048")
KEEP_GAIN = 0.015  # end-to-end gain an entry must show (run-to-run spread of the workloads is about 1%)


def clog2(v):
    b = 0
    while (1 << b) < v:
        b += 1
    return b


def bucket(m, n, k, kind):
    return (clog2(m), clog2(n), clog2(k), KIND_FLAGS[kind])


def workload_shapes(batch):
    """(workload, m, n, k, kind) of the ZC^T and FP64-solve GEMMs."""
    out = []
    for s in batch:
        out.append(("batch32", s.N, s.K, s.M, 1))
        out.append(("batch32", s.K, s.N, s.K, 0))
    for name in CONFIG_WORKLOADS:
        c = synthetic.CONFIGS[name]
        n = c["n"] ** 3
        m = c["M"]
        k = c["thc_cfac"] * m
        out.append((name, n, k, m, 1))
        out.append((name, k, n, k, 0))
    return out


def set_table(entries):
    flat = (ctypes.c_int * (5 * len(entries)))(*[int(v) for e in entries for v in e])
    L.check(L.lib().thc_gemm_tuning(flat if entries else None, len(entries)), "thc_gemm_tuning")


def entry(b):
    return b["key"] + [b["cfg"]]


def run_workload(ctx, name, data):
    """One run of a workload: mixed + FP64 factorisations of a config, or the
    batch (mixed on 4 streams + FP64 on 1); wall ms."""
    t0 = time.perf_counter()
    if name == "batch32":
        ctx.factorize_batched(data, precision=L.PREC_MIXED, max_streams=4)
        ctx.factorize_batched(data, precision=L.PREC_FP64, max_streams=1)
    else:
        s, g, r, a, cg = data
        for prec in (L.PREC_MIXED, L.PREC_FP64):
            ctx.factorize(g, r, a, s.thc_cfac, s.max_it, s.n, cg, precision=prec)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3


def time_pair(ctx, workloads, data, table_off, table_on, reps):
    """Best-of-reps wall ms of the workloads without and with a candidate
    entry, the two tables alternating run by run so that drift hits both."""
    best = {0: float("inf"), 1: float("inf")}
    for r in range(reps + 1):
        for side, table in ((0, table_off), (1, table_on)):
            set_table(table)
            t = sum(run_workload(ctx, wl, data[wl]) for wl in workloads)
            if r > 0:  # the first round warms up both
                best[side] = min(best[side], t)
    return best[0], best[1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TUNING_FILE)
    ap.add_argument("--reps", type=int, default=3, help="timed launches per GEMM configuration")
    ap.add_argument("--wreps", type=int, default=5, help="timed runs per workload in the validation")
    a = ap.parse_args()
    ctx = Context(0)
    set_table([])  # measure against the built-in rule
    t0 = time.perf_counter()
    batch = synthetic.make_batch(count=32, seed=0, max_it=20)
    shapes = workload_shapes(batch)
    print("autotune: %d GEMM shapes from %d workloads" % (len(shapes), 1 + len(CONFIG_WORKLOADS)), flush=True)

    # 1. every configuration on every distinct shape
    times = {}
    for (_, m, n, k, kind) in shapes:
        key = (m, n, k, kind)
        if key in times:
            continue
        row = {"rule": ctx.gemm_time(m, n, k, kind, 0, a.reps)}
        for cfg, nm in CFG_NAMES.items():
            row[nm] = ctx.gemm_time(m, n, k, kind, cfg, a.reps)
        times[key] = row
    groups = {}
    for (wl, m, n, k, kind) in shapes:
        groups.setdefault(bucket(m, n, k, kind), {"workloads": set(), "shapes": set()})
        g = groups[bucket(m, n, k, kind)]
        g["workloads"].add(wl)
        g["shapes"].add((m, n, k, kind))
    buckets = []
    for key, g in sorted(groups.items()):
        tot = {}
        for nm in ["rule"] + list(CFG_NAMES.values()):
            vals = [times[s][nm] for s in g["shapes"]]
            if all(v is not None for v in vals):
                tot[nm] = float(sum(vals))
        best = min((v, nm) for nm, v in tot.items() if nm != "rule")[1]
        cfg = [c for c, nm in CFG_NAMES.items() if nm == best][0]
        # an entry only where a configuration beats the rule by more than 1%
        gain = tot[best] < 0.99 * tot["rule"]
        buckets.append({"key": list(key), "workloads": sorted(g["workloads"]),
                        "shapes": [list(s) for s in sorted(g["shapes"])],
                        "ms": {nm: round(v, 4) for nm, v in tot.items()}, "best": best, "cfg": cfg,
                        "beats_rule": bool(gain)})
        print("bucket %s: best %s %s" % (key, best, {nm: round(v, 3) for nm, v in tot.items()}), flush=True)

    # 2. end-to-end validation, one candidate entry at a time (largest measured
    #    gain first): the workloads that have shapes in its bucket are timed
    #    with the entries kept so far, with and without it (alternating), and it
    #    is kept only if their summed time drops by more than KEEP_GAIN
    data = {"batch32": [{"grid": ctx.f64(s.grid), "rho": ctx.f64(s.rho), "aoR": ctx.f64(s.aoR),
                         "coulG": ctx.f64(s.coulG), "fft_grid": s.n, "thc_cfac": s.thc_cfac, "max_it": s.max_it,
                         "seed": s.seed} for s in batch]}
    for wl in CONFIG_WORKLOADS:
        s = synthetic.make_config(wl)
        data[wl] = (s, ctx.f64(s.grid), ctx.f64(s.rho), ctx.f64(s.aoR), ctx.f64(s.coulG))
    kept = set()
    validation = []
    cands = sorted((b for b in buckets if b["beats_rule"]), key=lambda b: b["ms"][b["best"]] - b["ms"]["rule"])
    for b in cands:
        base_entries = [entry(c) for c in buckets if id(c) in kept]
        off, on = time_pair(ctx, b["workloads"], data, base_entries, base_entries + [entry(b)], a.wreps)
        keep = on < (1.0 - KEEP_GAIN) * off
        if keep:
            kept.add(id(b))
        validation.append({"key": b["key"], "cfg": b["best"], "workloads": b["workloads"],
                           "ms_without": round(off, 3), "ms_with": round(on, 3), "kept": bool(keep)})
        print("validate %s -> %s on %s: %.2f -> %.2f ms: %s" % (b["key"], b["best"], ",".join(b["workloads"]), off,
                                                                on, "kept" if keep else "dropped"), flush=True)
    del data
    entries = [entry(b) for b in buckets if id(b) in kept]
    for b in buckets:
        b["kept"] = id(b) in kept
    out = {"arch": "gfx950", "device": torch.cuda.get_device_name(0), "command": "python -m thcpp_amd.autotune",
           "entry_format": "[ceil log2 m, ceil log2 n, ceil log2 k, flags, config]; flags 1 gathered | 2 lower | "
                           "4 square epilogue | 8 triangular; config 1 tiny 2 small 3 large 4 dtl 5 deepk",
           "entries": entries, "buckets": buckets, "validation": validation,
           "seconds": round(time.perf_counter() - t0, 1)}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    set_table(entries)
    print("autotune: %d entries kept, written to %s" % (len(entries), a.out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
