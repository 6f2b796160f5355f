"""thc_cholesky_noalloc (the no-allocation variant of thc_cholesky, for callers
that cannot allocate device memory) through the C ABI: the same factor and
info, bit for bit, as thc_cholesky, on two calls in a row with the second
smaller, and on a matrix whose leading minor fails."""
import ctypes

import numpy as np
import pytest
import torch

from thcpp_amd import _lib as L

pytestmark = pytest.mark.gpu


def _spd(k, seed):
    b = np.random.default_rng(seed).standard_normal((k, k + 7))
    return b @ b.T / k + 0.05 * np.eye(k)


def _noalloc(ctx, a):
    m = ctx.f64(a).clone()
    info = ctypes.c_int(-1)
    st = L.lib().thc_cholesky_noalloc(ctx.h, ctypes.c_void_p(m.data_ptr()), m.shape[0], ctypes.byref(info))
    return st, m, info.value


def test_noalloc_cholesky_matches_allocating_entry_point(ctx):
    for k in (777, 130):  # the second call is smaller and reuses the scratch
        a = _spd(k, k)
        ref, info_ref = ctx.cholesky(a)
        st, got, info = _noalloc(ctx, a)
        assert st == L.THC_OK and info == info_ref == 0
        assert torch.equal(got, ref)


def test_noalloc_cholesky_reports_failing_minor(ctx):
    """The info word is cleared at every call: a failing matrix, then a good
    one, report 1-based minor then 0."""
    k = 200
    a = np.eye(k)
    a[63, 63] = -1.0
    ref, info_ref = ctx.cholesky(a)
    st, got, info = _noalloc(ctx, a)
    assert st == L.THC_OK and info == info_ref == 64
    st, got, info = _noalloc(ctx, _spd(96, 1))
    assert st == L.THC_OK and info == 0
    ref, _ = ctx.cholesky(_spd(96, 1))
    assert torch.equal(got, ref)


def test_noalloc_cholesky_rejects_bad_arguments(ctx):
    info = ctypes.c_int(0)
    assert L.lib().thc_cholesky_noalloc(ctx.h, None, 10, ctypes.byref(info)) == L.THC_ERR_ARG
