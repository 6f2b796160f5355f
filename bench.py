"""ISDF factorisation throughput on one MI355X.

One step = one complete factorisation of one system with inputs resident in
HBM: weighted k-means (T iterations + nearest-grid mapping), ZC^T, CC^T,
Cholesky + inverse, the least-squares solve, the Coulomb-weighted FFT, V and
Luv = chol(V).  The workload is the 2x2x2 diamond supercell (64 atoms,
64^3 grid, 128 orbitals, N_mu = 2048, 30 k-means iterations) with synthetic
orbitals; the headline value is the mixed-precision path, the plain FP64 path
is timed beside it on the same inputs.

The CPU leg runs the oracle (the FP64 CPU restatement of the reference) over
one complete factorisation of the same system and is both the CPU baseline
and the reference the GPU results are checked against: k-means points
(identical), and the relative error of xi and V of every GPU path.

    python bench.py [--steps K] [--warmup W] [--config NAME] [--no-cpu] [--no-batch]
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

PEAK = {"hbm_gbs": 8000.0, "fp64_tflops": 78.6, "fp32_tflops": 157.3, "bf16_tflops": 2500.0, "int8_tops": 5000.0}
SV_DIGITS = 7                        # digits of the int8 solve; products with p + q < 7
SV_PRODUCTS = SV_DIGITS * (SV_DIGITS + 1) // 2
OZ_V_DIGITS = 5                      # digits per value of the int8 V (written by FFT pass B)
# HBM bytes per launch of the mixed path's solve GEMM (sv_gemm) on diamond64_n64_k2048, from rocprofv3
# FETCH_SIZE / WRITE_SIZE counters scaled by the same counters on a 1 GiB device copy (fetch x 2.0,
# write x 1.0); the summary with its commands is profiles/isdf_diamond64_n64_k2048_hbm_traffic.txt.
# Other workloads report null.
SOLVE_TRAFFIC_BYTES = {"diamond64_n64_k2048": 18916036 * 1024 * 2.0 + 2097152 * 1024 * 1.0}
METRIC = "ISDF factorizations/s + % roofline; mixed vs FP64 GPU at <=1e-6 rel V error"
STAGES = ["zct", "cct_chol_inv", "solve", "fft", "V", "luv", "isdf_total", "refine", "solve_gemm_kernel",
          "V_kernel", "fft_pass_a", "fft_pass_b"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="diamond64_n64_k2048")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg (and the errors against it)")
    p.add_argument("--no-batch", action="store_true", help="skip the 32-cell batched leg and the 8-atom cell leg")
    p.add_argument("--no-variants", action="store_true", help="skip the FP32 and refinement legs")
    p.add_argument("--batch-only", action="store_true",
                   help="run only the mixed 32-cell batch on 4 streams (--steps batches after --warmup)")
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


def cpu_oracle(s):
    """One complete factorisation by the oracle on the host's cores: k-means
    (classify on OpenMP threads, the rest sequential, as the reference at one
    rank), ZC^T, CC^T, the QR least-squares solve, the forward FFTs, Muv and
    its Cholesky, with the multithreaded BLAS/FFT numpy and scipy bring."""
    from oracle import isdf_oracle as IO
    from oracle import kmeans_oracle as KO
    cores = len(os.sched_getaffinity(0))
    t = {}
    t0 = time.perf_counter()
    km = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0)
    t["kmeans"] = time.perf_counter() - t0
    idx = km["interp_idx"]
    t0 = time.perf_counter()
    z = IO.czt(s.aoR, idx)
    c = IO.cct(z, idx)
    t["zct_cct"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    x = IO.least_squares(c, z)
    t["solve"] = time.perf_counter() - t0
    del z
    t0 = time.perf_counter()
    ivg = IO.fft_vectors(x, s.n, workers=cores)
    t["fft"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    v = IO.construct_muv(ivg, s.coulG, real_part=True)
    t["muv"] = time.perf_counter() - t0
    del ivg
    t0 = time.perf_counter()
    IO.cholesky_luv(v)
    t["luv"] = time.perf_counter() - t0
    total = sum(t.values())
    base = {"value": round(1.0 / total, 5), "unit": "factorizations/s", "cores": cores, "kind": "port",
            "sample": "one complete factorisation of the same system, no extrapolation; stage seconds " +
                      " ".join("%s %.3f" % kv for kv in t.items())}
    return base, {"idx": idx, "X": x, "V": v}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_rows(dev, host):
    """||dev - host|| / ||host|| over a K x N array kept on the device, row block by row block."""
    import torch
    num = den = 0.0
    for r0 in range(0, host.shape[0], 256):
        h = torch.from_numpy(np.ascontiguousarray(host[r0:r0 + 256]))
        d = dev[r0:r0 + 256].double().cpu()
        num += float(torch.sum((d - h) ** 2))
        den += float(torch.sum(h ** 2))
    return (num / den) ** 0.5


def batch_systems(ctx):
    """The uneven 32-cell batch, its inputs resident on the GPU."""
    from thcpp_amd import synthetic
    batch = synthetic.make_batch(count=32, seed=0, max_it=20)
    systems = [{"grid": ctx.f64(s.grid), "rho": ctx.f64(s.rho), "aoR": ctx.f64(s.aoR), "coulG": ctx.f64(s.coulG),
                "fft_grid": s.n, "thc_cfac": s.thc_cfac, "max_it": s.max_it, "seed": s.seed} for s in batch]
    return batch, systems


def batch_only(steps, warmup):
    """Only the mixed-precision 32-cell batch on 4 streams (for profiling the
    batched path by itself): one JSON line with its time per batch."""
    import torch
    from thcpp_amd import _lib as L
    from thcpp_amd.api import Context
    ctx = Context(0)
    batch, systems = batch_systems(ctx)
    for _ in range(warmup):
        ctx.factorize_batched(systems, precision=L.PREC_MIXED, max_streams=4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st, _res = ctx.factorize_batched(systems, precision=L.PREC_MIXED, max_streams=4)
        if st != L.THC_OK:
            raise RuntimeError("batched factorisation status %d" % st)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"workload": "batch of 32 uneven diamond cells, 20 k-means iterations each, mixed, 4 streams",
                      "steps": steps, "warmup": warmup, "ms_per_batch": round(dt * 1e3, 3),
                      "systems_per_s": round(len(batch) / dt, 2)}))
    return 0


def batched_leg(ctx, steps, warmup):
    """The uneven 32-cell batch (side 6-14 bohr, 8-64 atoms, grids 24^3-48^3,
    N_mu = 8-10 M) through thc_factorize_batched, mixed precision, on 4
    streams and, for comparison, on 1, and in FP64 on 1; then the plain FP64
    GPU path -- thc_factorize one system after the other, no batching -- the
    baseline of the batched speed-up."""
    import torch
    from thcpp_amd import _lib as L
    batch, systems = batch_systems(ctx)
    out = {"workload": "batch of 32 uneven diamond cells, 20 k-means iterations each",
           "systems": len(batch), "sum_N": int(sum(s.N for s in batch)),
           "sum_N_mu": int(sum(s.K for s in batch))}
    for name, prec, streams in (("streams_4", L.PREC_MIXED, 4), ("streams_1", L.PREC_MIXED, 1),
                                ("fp64_streams_1", L.PREC_FP64, 1)):
        for _ in range(warmup):
            ctx.factorize_batched(systems, precision=prec, max_streams=streams)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st, _res = ctx.factorize_batched(systems, precision=prec, max_streams=streams)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        if st != L.THC_OK:
            raise RuntimeError("batched factorisation status %d" % st)
        out[name] = {"ms_per_batch": round(dt * 1e3, 3), "systems_per_s": round(len(batch) / dt, 2)}
    cells = [(ctx.f64(s.grid), ctx.f64(s.rho), ctx.f64(s.aoR), ctx.f64(s.coulG), s) for s in batch]

    def one_at_a_time():
        for g, r, a, cg, s in cells:
            o = ctx.factorize(g, r, a, s.thc_cfac, s.max_it, s.n, cg, seed=s.seed, precision=L.PREC_FP64)
            if o["status"] != L.THC_OK:
                raise RuntimeError("factorisation status %d" % o["status"])
    for _ in range(warmup):
        one_at_a_time()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_at_a_time()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    out["fp64_one_at_a_time"] = {"ms_per_batch": round(dt * 1e3, 3), "systems_per_s": round(len(batch) / dt, 2)}
    out["mixed_streams_4_over_fp64_one_at_a_time"] = round(out["fp64_one_at_a_time"]["ms_per_batch"] /
                                                           out["streams_4"]["ms_per_batch"], 3)
    return out


def small_cell_leg(ctx, steps, with_cpu):
    """The 8-atom cell (32^3 grid, M = 16, N_mu = 128, 20 k-means iterations),
    the case the reference runs on a CPU: GPU FP64 and mixed factorisation
    rates and, with with_cpu, the CPU oracle's complete factorisation and the
    GPU results against it."""
    import torch
    from thcpp_amd import _lib as L
    from thcpp_amd import synthetic
    s = synthetic.make_config("diamond8_n32")
    g, r, ao, cg = ctx.f64(s.grid), ctx.f64(s.rho), ctx.f64(s.aoR), ctx.f64(s.coulG)
    out = {"workload": "diamond8_n32_k128", "N": s.N, "M": s.M, "N_mu": s.K, "kmeans_iters": s.max_it}
    res = {}
    for name, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
        for _ in range(3):
            o = ctx.factorize(g, r, ao, s.thc_cfac, s.max_it, s.n, cg, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            o = ctx.factorize(g, r, ao, s.thc_cfac, s.max_it, s.n, cg, precision=prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        if o["status"] != L.THC_OK:
            raise RuntimeError("factorisation status %d" % o["status"])
        res[name] = o
        out[name] = {"ms_per_step": round(dt * 1e3, 3), "value": round(1.0 / dt, 2)}
    if with_cpu:
        base, ref = cpu_oracle(s)
        out["cpu_baseline"] = base
        out["kmeans_points_identical"] = bool(all(np.array_equal(res[p]["interp_idx"].cpu().numpy(), ref["idx"])
                                                  for p in res))
        for name in res:
            out["rel_V_vs_oracle_" + name] = rel(res[name]["V"].double().cpu().numpy(), ref["V"])
    return out


def bandwidth(stages, s, x_bytes, es, digits=False):
    """Achieved GB/s of the HBM-bound stages from their device times: each FFT
    pass reads and writes the k-row block once (pass A: x in, the complex
    z/y-transformed half spectrum out; pass B: that spectrum and the weights
    in, the weighted spectrum out -- or, with digits, V's five int8 digit
    planes, each y-line of 2 n0 (n2/2+1) bytes padded to a multiple of 64,
    and one int32 exponent per (row, y-line)); k-means reads the coordinates
    (24 B) and the density (8 B) and writes the assignment (4 B) of every
    point once per iteration."""
    K, N = s.K, s.N
    nh = s.n[0] * s.n[1] * (s.n[2] // 2 + 1)
    a_bytes = K * N * x_bytes + K * nh * 2 * es
    if digits:
        lp = (2 * s.n[0] * (s.n[2] // 2 + 1) + 63) // 64 * 64
        b_bytes = K * nh * 2 * es + nh * es + OZ_V_DIGITS * K * s.n[1] * lp + 4 * K * s.n[1]
    else:
        b_bytes = 2 * K * nh * 2 * es + nh * es
    km_bytes = 36.0 * N * s.max_it
    out = {}
    for name, nbytes, ms in (("fft_pass_a", a_bytes, stages[10]), ("fft_pass_b", b_bytes, stages[11]),
                             ("kmeans", km_bytes, stages[12])):
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[name] = {"bytes": float(nbytes), "ms": round(float(ms), 3), "GB_s": round(gbs, 1),
                     "frac_hbm": round(gbs / PEAK["hbm_gbs"], 3)}
    return out


def main():
    a = parse()
    if a.copy_cal:
        return copy_calibration()
    if a.gpus != 1:
        print("bench.py measures one GPU only (the ISDF path has no multi-GPU sharding); run with --gpus 1",
              file=sys.stderr)
        return 2
    if a.batch_only:
        return batch_only(a.steps, a.warmup)
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

    def run(prec, refine=0):
        return ctx.factorize(g, r, ao, s.thc_cfac, s.max_it, s.n, cg, precision=prec, refine=refine)

    def timed(prec, refine=0, steps=a.steps, warmup=a.warmup):
        for _ in range(warmup):
            run(prec, refine)
        torch.cuda.synchronize()
        stages = []
        t0 = time.perf_counter()
        for _ in range(steps):
            out = run(prec, refine)
            stages.append(out["stage_ms"] + [out["kmeans_ms"]])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        if out["status"] != L.THC_OK:
            raise RuntimeError("factorisation status %d" % out["status"])
        print("bench: precision %d refine %d: %.3f ms per factorisation" % (prec, refine, dt * 1e3), file=sys.stderr,
              flush=True)
        return {"ms": dt * 1e3, "stages": np.mean(np.array(stages), axis=0), "V": out["V"].double().cpu().numpy(),
                "idx": out["interp_idx"].cpu().numpy()}

    res = {"fp64": timed(L.PREC_FP64), "mixed": timed(L.PREC_MIXED)}
    mx = res["mixed"]
    ms = mx["ms"]
    st = mx["stages"]  # STAGES order, then kmeans
    K, N = s.K, s.N
    gemm_ms = float(st[8])
    sv_ops = SV_PRODUCTS * 2.0 * K * K * N
    achieved = sv_ops / (gemm_ms * 1e-3) / 1e12
    kp = -(-K // 32) * 32
    kr, nr = -(-K // 128) * 128, -(-N // 64) * 64
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
                   "M": s.M, "N_mu": K, "kmeans_iters": s.max_it,
                   "precision": "mixed: FP64 k-means, ZC^T and Cholesky/inverse of CC^T; solve from exact int8 "
                                "MFMA products of 7-digit splits (28 products); FP32 FFT; V from exact int8 MFMA "
                                "products of 5 digits; FP64 sums",
                   "parallelism": "single GPU"},
        "fp64_path": {"value": round(1000.0 / res["fp64"]["ms"], 3), "ms_per_step": round(res["fp64"]["ms"], 3),
                      "stage_ms": {k: round(float(v), 3) for k, v in zip(STAGES + ["kmeans"], res["fp64"]["stages"])}},
        "mixed_over_fp64_speedup": round(res["fp64"]["ms"] / ms, 3),
        "bandwidth": {"mixed": bandwidth(st, s, 4, 4, digits=True), "fp64": bandwidth(res["fp64"]["stages"], s, 8, 8)},
        "rel_V_error_mixed_vs_fp64": rel(mx["V"], res["fp64"]["V"]),
        "stage_ms": {k: round(float(v), 3) for k, v in zip(STAGES + ["kmeans"], st)},
        "roofline": {"kernel": "solve GEMM X = CC^T^-1 ZC^T as exact int8 digit products (sv_gemm, "
                               "v_mfma_i32_32x32x32_i8); int8 multiply-adds counted as 2 ops",
                     "bound": "mfma",
                     "achieved": round(achieved, 1), "peak": PEAK["int8_tops"], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["int8_tops"], 4),
                     "traffic": SOLVE_TRAFFIC_BYTES.get(a.config),
                     "traffic_unit": "bytes per launch",
                     # digit planes read once (7 planes of the padded K x K and N x K operands), X written in FP32
                     "algorithmic_bytes": float(SV_DIGITS * (kr * kp + nr * kp) + 4.0 * K * N),
                     "algorithmic_ops": sv_ops, "kernel_ms": round(gemm_ms, 3)},
    }
    if not a.no_variants:
        f32 = timed(L.PREC_FP32)
        ref1 = timed(L.PREC_MIXED, refine=1, steps=2, warmup=1)
        line["variants"] = {
            "fp32": {"ms_per_step": round(f32["ms"], 3), "rel_V_vs_fp64_path": rel(f32["V"], res["fp64"]["V"])},
            "mixed_refine1": {"ms_per_step": round(ref1["ms"], 3),
                              "rel_V_vs_fp64_path": rel(ref1["V"], res["fp64"]["V"])}}
    if not a.no_batch:
        line["batched"] = batched_leg(ctx, max(1, min(a.steps, 3)), 1)
        line["small_cell"] = small_cell_leg(ctx, 20, not a.no_cpu)
    if not a.no_cpu:
        base, ref = cpu_oracle(s)
        line["cpu_baseline"] = base
        acc = {"kmeans_points_identical": bool(all(np.array_equal(res[p]["idx"], ref["idx"]) for p in res))}
        for name, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
            out = ctx.isdf(ao, ref["idx"], s.n, cg, precision=prec, want_xi=True)
            acc["rel_xi_vs_oracle_" + name] = rel_rows(out["xi"], ref["X"])
            acc["rel_V_vs_oracle_" + name] = rel(out["V"].double().cpu().numpy(), ref["V"])
            del out
        if "variants" in line:
            out = ctx.isdf(ao, ref["idx"], s.n, cg, precision=L.PREC_FP32)
            acc["rel_V_vs_oracle_fp32"] = rel(out["V"].double().cpu().numpy(), ref["V"])
        line["accuracy_vs_oracle"] = acc
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
