"""Batched factorisation of uneven cells (thc_factorize_batched): every
system's result equals the single-system factorisation on the same inputs,
its k-means points equal the oracle's (src/kmeans.cpp:187-264, one system at a
time as src/main.cpp:50-60 runs it), and a failing system leaves the others
computed."""
import numpy as np
import pytest
import torch

from oracle import kmeans_oracle as KO
from thcpp_amd import _lib as L
from thcpp_amd import synthetic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    # grids 24^3 .. 45^3 (odd and 3/5-smooth edges), M 16..80, N_mu 128..720
    return synthetic.make_batch(count=6, seed=3, max_it=6)


def as_dict(s):
    return {"grid": s.grid, "rho": s.rho, "aoR": s.aoR, "coulG": s.coulG, "fft_grid": s.n,
            "thc_cfac": s.thc_cfac, "max_it": s.max_it, "seed": s.seed}


@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_MIXED])
@pytest.mark.parametrize("streams", [1, 4])
def test_batched_equals_single(ctx, batch, prec, streams):
    print("batched vs single: precision %d, %d streams, %d systems" % (prec, streams, len(batch)), flush=True)
    st, outs = ctx.factorize_batched([as_dict(s) for s in batch], precision=prec, max_streams=streams)
    assert st == L.THC_OK, (st, [o["status"] for o in outs], L.lib().thc_last_error())
    for s, o in zip(batch, outs):
        assert o["status"] == L.THC_OK
        assert o["iters"] == s.max_it
        one = ctx.factorize(s.grid, s.rho, s.aoR, s.thc_cfac, s.max_it, s.n, s.coulG, precision=prec)
        # same kernels on the same shapes, fixed reduction orders: bitwise equal
        assert torch.equal(o["interp_idx"], one["interp_idx"])
        assert torch.equal(o["V"], one["V"])
        assert torch.equal(o["Luv"], one["Luv"])


def test_batched_kmeans_matches_oracle(ctx, batch):
    st, outs = ctx.factorize_batched([as_dict(s) for s in batch], precision=L.PREC_MIXED)
    assert st == L.THC_OK
    for s, o in zip(batch, outs):
        ref = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, s.seed)
        assert ref["status"] == 0
        assert np.array_equal(o["interp_idx"].cpu().numpy(), ref["interp_idx"]), s.n


@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_MIXED])
def test_batched_v_matches_oracle(ctx, batch, prec):
    """V of three cells of the batch (smallest, middle, largest grid) against
    the oracle's FP64 ISDF on the oracle's k-means points: FP64 within the
    conditioning floor, mixed within 1e-6 beyond it."""
    from oracle import isdf_oracle as IO
    st, outs = ctx.factorize_batched([as_dict(s) for s in batch], precision=prec)
    assert st == L.THC_OK
    order = np.argsort([s.N for s in batch])
    for i in (order[0], order[len(order) // 2], order[-1]):
        s, o = batch[i], outs[i]
        idx = KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, s.seed)["interp_idx"]
        v_ref = IO.isdf(s.aoR, idx, s.n, s.coulG)["V"]
        fp64 = ctx.isdf(s.aoR, idx, s.n, s.coulG)
        e64 = np.linalg.norm(fp64["V"].cpu().numpy() - v_ref) / np.linalg.norm(v_ref)
        e = np.linalg.norm(o["V"].cpu().numpy() - v_ref) / np.linalg.norm(v_ref)
        print("cell %d (grid %s, K %d) prec %d: rel V vs oracle %.3e (FP64 path %.3e)" % (i, s.n, s.K, prec, e, e64))
        assert e64 < 1e-6
        assert e < e64 + (1e-6 if prec == L.PREC_MIXED else 1e-12)


@pytest.mark.parametrize("streams", [4, 1])
def test_batched_error_isolated(ctx, batch, streams):
    """A system with thc_cfac = 0 (no interpolation points) fails with
    THC_ERR_ARG; the rest are still factorised and the call returns that
    system's status.  On one stream the failing system is in the second
    k-means run (all but the largest system), which goes on beside the
    largest system's ISDF stages; on four streams (three systems) there is
    no second run."""
    systems = [as_dict(s) for s in batch[:3]]
    bad = dict(systems[1])
    bad["thc_cfac"] = 0
    systems[1] = bad
    st, outs = ctx.factorize_batched(systems, precision=L.PREC_FP64, max_streams=streams)
    assert st == L.THC_ERR_ARG
    assert outs[1]["status"] == L.THC_ERR_ARG
    for i in (0, 2):
        s = batch[i]
        assert outs[i]["status"] == L.THC_OK
        one = ctx.factorize(s.grid, s.rho, s.aoR, s.thc_cfac, s.max_it, s.n, s.coulG)
        assert torch.equal(outs[i]["V"], one["V"])


def test_batched_isdf_error_isolated(ctx, batch):
    """A system whose FFT grid does not hold its N points passes the batched
    k-means and fails in its ISDF stages on a lane (THC_ERR_ARG); the lanes go
    on taking the other systems, which equal their single factorisations."""
    systems = [as_dict(s) for s in batch[:4]]
    bad = dict(systems[2])
    n = tuple(bad["fft_grid"])
    bad["fft_grid"] = (n[0], n[1], n[2] + 1)
    systems[2] = bad
    st, outs = ctx.factorize_batched(systems, precision=L.PREC_MIXED, max_streams=4)
    assert st == L.THC_ERR_ARG
    assert outs[2]["status"] == L.THC_ERR_ARG
    for i in (0, 1, 3):
        s = batch[i]
        assert outs[i]["status"] == L.THC_OK
        one = ctx.factorize(s.grid, s.rho, s.aoR, s.thc_cfac, s.max_it, s.n, s.coulG, precision=L.PREC_MIXED)
        assert torch.equal(outs[i]["V"], one["V"])


def test_batched_empty(ctx):
    st, outs = ctx.factorize_batched([])
    assert st == L.THC_OK and outs == []


def test_batched_kmeans_uneven_loops(ctx, batch):
    """Systems with different iteration counts, seeds and a threshold that
    stops one loop after its first iteration, run in one batched k-means:
    each system's points, iteration count and V equal its own factorisation."""
    systems = []
    for i, (max_it, thr, seed) in enumerate([(3, -1.0, 0), (7, -1.0, 5), (12, -1.0, 2), (6, 1e10, 0)]):
        d = as_dict(batch[i])
        d["max_it"], d["threshold"], d["seed"] = max_it, thr, seed
        systems.append(d)
    print("batched k-means, uneven loops: %d systems" % len(systems), flush=True)
    st, outs = ctx.factorize_batched(systems, precision=L.PREC_FP64, max_streams=2)
    assert st == L.THC_OK, (st, [o["status"] for o in outs], L.lib().thc_last_error())
    for s, d, o in zip(batch, systems, outs):
        one = ctx.factorize(s.grid, s.rho, s.aoR, s.thc_cfac, d["max_it"], s.n, s.coulG, threshold=d["threshold"],
                            seed=d["seed"])
        assert o["iters"] == (1 if d["threshold"] > 0 else d["max_it"])
        assert torch.equal(o["interp_idx"], one["interp_idx"])
        assert torch.equal(o["V"], one["V"])
