"""ISDF factorisation throughput on one MI355X.

One step = one complete factorisation of one system with inputs resident in
HBM: weighted k-means (T iterations + nearest-grid mapping), ZC^T, CC^T,
Cholesky + inverse, the least-squares solve, the Coulomb-weighted FFT, V and
Luv = chol(V).  The workload is the 2x2x2 diamond supercell (64 atoms,
64^3 grid, 128 orbitals, N_mu = 2048, 30 k-means iterations) with synthetic
orbitals; the headline value is the mixed-precision path, the plain FP64 path
is timed beside it on the same inputs, and the relative V error between the
two is reported.

    python bench.py [--steps K] [--warmup W] [--config NAME] [--no-cpu]
    python bench.py --copy-cal     # 1 GiB device copy x3, for HBM counter calibration
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm_gbs": 8000.0, "fp64_tflops": 78.6, "fp32_tflops": 157.3, "bf16_tflops": 2500.0}
# HBM bytes per launch of the mixed-path solve GEMM on diamond64_n64_k2048, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE counters scaled by the same counters on a 1 GiB device copy (fetch x 2.0, write x 1.0); the
# summary with its commands is profiles/isdf_diamond64_n64_k2048_hbm_traffic.txt.  Other workloads
# report null.
SOLVE_TRAFFIC_BYTES = {"diamond64_n64_k2048": 38.42e9}
METRIC = "ISDF factorizations/s + % roofline; mixed vs FP64 GPU at <=1e-6 rel V error"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="diamond64_n64_k2048")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-batch", action="store_true", help="skip the 32-cell batched leg")
    p.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU sample work")
    p.add_argument("--copy-cal", action="store_true",
                   help="only run a 1 GiB float64 device copy three times (the known-size kernel that "
                        "FETCH_SIZE / WRITE_SIZE counters are scaled against) and exit")
    return p.parse_args()


def copy_calibration():
    import torch
    n = 1 << 27  # 2^27 doubles = 1 GiB
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
    return 0


def cpu_baseline(s, budget):
    """The oracle (the CPU FP64 restatement of the reference) on a bounded
    sample of the same workload, scaled to one factorisation:
    k-means: one sequential classify + update pass over the full grid, times T+1
    (plus one map_to_grid pass, same cost); ISDF: ZC^T, the QR solve, the FFT and
    Muv on a column slice of the grid / a subset of mu, scaled linearly."""
    import scipy.linalg
    from oracle import isdf_oracle as IO
    from oracle import kmeans_oracle as KO
    cores = len(os.sched_getaffinity(0))
    idx0 = KO.initial_indices(s.N, s.K, 0)
    cen = s.grid[idx0]
    t = time.perf_counter()
    gmap = KO.classify(s.grid, cen)
    KO.update_centroids(s.rho, s.grid, gmap, s.K)
    t_iter = time.perf_counter() - t
    t_kmeans = t_iter * (s.max_it + 2)
    # ISDF pieces on a fraction f of the grid columns / of the mu rows
    f = 1.0 / 16
    idx = np.sort(idx0)
    ncols = int(s.N * f)
    t = time.perf_counter()
    z_full_idx = IO.czt(s.aoR, idx)  # needed whole for CC^T (Z[:, idx])
    t_z = time.perf_counter() - t
    c = IO.cct(z_full_idx, idx)
    t = time.perf_counter()
    q, r = scipy.linalg.qr(c, mode="economic")
    t_qr = time.perf_counter() - t
    t = time.perf_counter()
    x_part = scipy.linalg.solve_triangular(r, q.T @ z_full_idx[:, :ncols], lower=False)
    t_apply = (time.perf_counter() - t) / f
    del x_part
    kmu = max(1, int(s.K * f))
    x = z_full_idx[:kmu]
    t = time.perf_counter()
    ivg = IO.fft_vectors(x, s.n)
    t_fft = (time.perf_counter() - t) * s.K / kmu
    a = ivg[:ncols]
    t = time.perf_counter()
    _ = a.conj().T @ a
    t_muv = (time.perf_counter() - t) / f * (s.K / kmu) ** 2
    total = t_kmeans + t_z + t_qr + t_apply + t_fft + t_muv
    return {"value": 1.0 / total, "unit": "factorizations/s", "cores": cores, "kind": "port",
            "sample": ("oracle FP64 restatement on the same system: 1 k-means pass x (T+2), ZC^T in full, "
                       "QR of CC^T, QR solve on 1/16 of the grid columns, FFT of %d of %d vectors, Muv on "
                       "1/16 of G for those vectors, each scaled to a full factorisation; stage seconds "
                       "kmeans %.1f zct %.1f qr %.1f solve %.1f fft %.1f muv %.1f"
                       % (kmu, s.K, t_kmeans, t_z, t_qr, t_apply, t_fft, t_muv))}


def batched_leg(ctx, steps, warmup):
    """The uneven 32-cell batch (side 6-14 bohr, 8-64 atoms, grids 24^3-48^3,
    N_mu = 8-10 M) through thc_factorize_batched, mixed precision, on 4
    streams and, for comparison, on 1."""
    import torch
    from thcpp_amd import _lib as L
    from thcpp_amd import synthetic
    batch = synthetic.make_batch(count=32, seed=0, max_it=20)
    systems = [{"grid": ctx.f64(s.grid), "rho": ctx.f64(s.rho), "aoR": ctx.f64(s.aoR), "coulG": ctx.f64(s.coulG),
                "fft_grid": s.n, "thc_cfac": s.thc_cfac, "max_it": s.max_it, "seed": s.seed} for s in batch]
    out = {"workload": "batch of 32 uneven diamond cells, 20 k-means iterations each",
           "systems": len(batch), "sum_N": int(sum(s.N for s in batch)),
           "sum_N_mu": int(sum(s.K for s in batch))}
    for streams in (4, 1):
        for _ in range(warmup):
            ctx.factorize_batched(systems, precision=L.PREC_MIXED, max_streams=streams)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st, _res = ctx.factorize_batched(systems, precision=L.PREC_MIXED, max_streams=streams)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        if st != L.THC_OK:
            raise RuntimeError("batched factorisation status %d" % st)
        out["streams_%d" % streams] = {"ms_per_batch": round(dt * 1e3, 3),
                                       "systems_per_s": round(len(batch) / dt, 2)}
    return out


def main():
    a = parse()
    if a.copy_cal:
        return copy_calibration()
    if a.gpus != 1:
        print("bench.py measures one GPU only (the ISDF path has no multi-GPU sharding); run with --gpus 1",
              file=sys.stderr)
        return 2
    import torch
    from thcpp_amd import _lib as L
    from thcpp_amd import synthetic
    from thcpp_amd.api import Context

    t0 = time.perf_counter()
    s = synthetic.make_config(a.config)
    print("bench: %s generated in %.1f s (N=%d, M=%d, N_mu=%d)" % (a.config, time.perf_counter() - t0, s.N, s.M, s.K),
          file=sys.stderr, flush=True)
    ctx = Context(0)
    g, r, ao, cg = ctx.f64(s.grid), ctx.f64(s.rho), ctx.f64(s.aoR), ctx.f64(s.coulG)

    def run(prec):
        return ctx.factorize(g, r, ao, s.thc_cfac, s.max_it, s.n, cg, precision=prec)

    results = {}
    for name, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
        for _ in range(a.warmup):
            run(prec)
        torch.cuda.synchronize()
        stages = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            out = run(prec)
            stages.append(out["stage_ms"] + [out["kmeans_ms"]])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        if out["status"] not in (L.THC_OK,):
            print("factorisation status %d" % out["status"], file=sys.stderr)
            return 1
        results[name] = {"ms": dt * 1e3, "stages": np.mean(np.array(stages), axis=0), "V": out["V"].double().cpu()}
        print("bench: %s path %.3f ms per factorisation" % (name, dt * 1e3), file=sys.stderr, flush=True)
    v64, vmx = results["fp64"]["V"], results["mixed"]["V"]
    relv = float(torch.linalg.norm(vmx - v64) / torch.linalg.norm(v64))
    mx = results["mixed"]
    ms = mx["ms"]
    st = mx["stages"]  # 0 zct,1 cct+chol+inv,2 solve,3 fft,4 V,5 luv,6 isdf total,7 refine,8 kmeans
    K, N = s.K, s.N
    solve_ms = float(st[2])
    solve_flops = 2.0 * K * K * N
    achieved = solve_flops / (solve_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(1000.0 / ms, 3),
        "unit": "factorizations/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32+i8",
        "data": "synthetic",
        "config": {"workload": a.config, "atoms": int(s.atoms.shape[0]), "grid": list(s.n), "N": N,
                   "M": s.M, "N_mu": K, "kmeans_iters": s.max_it, "precision": "mixed: FP64 k-means, ZC^T, "
                   "Cholesky, solve; FP32 FFT; V from exact int8 MFMA products of 5 digits, FP64 sum",
                   "parallelism": "single GPU"},
        "fp64_path": {"value": round(1000.0 / results["fp64"]["ms"], 3), "ms_per_step": round(results["fp64"]["ms"], 3)},
        "mixed_over_fp64_speedup": round(results["fp64"]["ms"] / ms, 3),
        "rel_V_error_mixed_vs_fp64": relv,
        "stage_ms": {k: round(float(v), 3) for k, v in zip(
            ["zct", "cct_chol_inv", "solve", "fft", "V", "luv", "isdf_total", "refine", "kmeans"], st)},
        "roofline": {"kernel": "solve GEMM X = CC^T^-1 ZC^T (gemm_dtl_kernel, FP64 MFMA)", "bound": "mfma",
                     "achieved": round(achieved, 2), "peak": PEAK["fp64_tflops"], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["fp64_tflops"], 4),
                     "traffic": SOLVE_TRAFFIC_BYTES.get(a.config),
                     "traffic_unit": "bytes per launch",
                     # Zt and (CC^T)^-1 read in FP64, X written in FP32 (nothing else reads X)
                     "algorithmic_bytes": 8.0 * (K * N + K * K) + 4.0 * K * N},
    }
    if not a.no_batch:
        line["batched"] = batched_leg(ctx, max(1, min(a.steps, 3)), 1)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(s, a.cpu_budget)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
