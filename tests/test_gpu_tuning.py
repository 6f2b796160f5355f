"""GEMM tile tuning: every configuration the tuning table can select gives the
ISDF result of the built-in rule (ZC^T and the FP64 solve go through it), and
the timing hook of the autotuner refuses a configuration that does not apply."""
import ctypes

import numpy as np
import pytest
import torch

from thcpp_amd import _lib as L
from thcpp_amd import synthetic
from thcpp_amd.api import apply_tuning
from thcpp_amd.autotune import bucket, set_table

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def small_system():
    grid = (16, 16, 15)
    n = int(np.prod(grid))
    rng = np.random.default_rng(4)
    a, _ = np.linalg.qr(rng.standard_normal((n, 12)))
    idx = np.sort(rng.choice(n, 60, replace=False)).astype(np.int32)
    return a, idx, grid, synthetic.coulomb_kernel(7.0, grid)


@pytest.mark.parametrize("cfg", [1, 2, 3, 4, 5])
def test_forced_config_isdf_matches_rule(ctx, small_system, cfg):
    a, idx, grid, cg = small_system
    n, m, k = a.shape[0], a.shape[1], idx.size
    print("tuned GEMM configuration %d" % cfg, flush=True)
    try:
        set_table([])
        ref = ctx.isdf(a, idx, grid, cg, want_xi=True)
        set_table([list(bucket(n, k, m, 1)) + [cfg], list(bucket(k, n, k, 0)) + [cfg]])
        out = ctx.isdf(a, idx, grid, cg, want_xi=True)
    finally:
        apply_tuning()
    assert out["status"] == L.THC_OK
    assert rel(out["xi"].cpu(), ref["xi"].cpu()) < 1e-11
    assert rel(out["V"].cpu(), ref["V"].cpu()) < 1e-11


def test_gemm_time_hook(ctx):
    ms = ctx.gemm_time(4096, 256, 256, kind=0, cfg=0, reps=2)
    assert ms is not None and ms > 0
    # direct-to-LDS tiles need plain, ungathered operands: refused for the ZC^T kind
    assert ctx.gemm_time(4096, 256, 32, kind=1, cfg=L.GEMM_DTL, reps=1) is None
    for cfg in (1, 2, 3, 5):
        assert ctx.gemm_time(4096, 256, 32, kind=1, cfg=cfg, reps=1) > 0
    assert L.lib().thc_gemm_tuning(None, -1) == L.THC_ERR_ARG
    bad = (ctypes.c_int * 5)(12, 8, 8, 0, 9)
    assert L.lib().thc_gemm_tuning(bad, 1) == L.THC_ERR_ARG
