"""ORACLE - TEST INFRASTRUCTURE ONLY.

ctypes front end for oracle/kmeans_oracle.cpp, the CPU restatement of the
reference's KMeansSolver (src/kmeans.cpp:88-264).  Built by oracle/Makefile
(g++ -O2 -ffp-contract=off); build() compiles it on demand.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libkmeans_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U = ctypes.POINTER(ctypes.c_uint)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "libkmeans_oracle.so"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        l = ctypes.CDLL(_LIB_PATH)
        l.oracle_initial_indices.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_uint, _I]
        l.oracle_mt19937.argtypes = [ctypes.c_uint, ctypes.c_long, _U]
        l.oracle_classify.argtypes = [_D, ctypes.c_long, _D, ctypes.c_long, _I]
        l.oracle_update_centroids.argtypes = [_D, _D, ctypes.c_long, _I, ctypes.c_long, _D]
        l.oracle_update_centroids.restype = ctypes.c_int
        l.oracle_normed_difference.argtypes = [_D, _D, ctypes.c_long]
        l.oracle_normed_difference.restype = ctypes.c_double
        l.oracle_percentage_change.argtypes = [_I, _I, ctypes.c_long]
        l.oracle_percentage_change.restype = ctypes.c_double
        l.oracle_map_to_grid.argtypes = [_D, ctypes.c_long, _D, ctypes.c_long, _I]
        l.oracle_map_to_grid.restype = ctypes.c_int
        l.oracle_kmeans.argtypes = [_D, _D, ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_uint, _I, _I, _D, _D, _I, _D, _I]
        l.oracle_kmeans.restype = ctypes.c_int
        _lib = l
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def initial_indices(num_grid: int, num_interp: int, seed: int) -> np.ndarray:
    out = np.empty(num_interp, dtype=np.int32)
    lib().oracle_initial_indices(num_grid, num_interp, seed, _ip(out))
    return out


def mt19937(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    lib().oracle_mt19937(seed, count, out.ctypes.data_as(_U))
    return out


def classify(grid: np.ndarray, centroids: np.ndarray) -> np.ndarray:
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    centroids = np.ascontiguousarray(centroids, dtype=np.float64)
    out = np.empty(grid.shape[0], dtype=np.int32)
    lib().oracle_classify(_dp(grid), grid.shape[0], _dp(centroids), centroids.shape[0], _ip(out))
    return out


def update_centroids(rho, grid, grid_map, k):
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    grid_map = np.ascontiguousarray(grid_map, dtype=np.int32)
    out = np.empty((k, 3))
    crit = lib().oracle_update_centroids(_dp(rho), _dp(grid), grid.shape[0], _ip(grid_map), k, _dp(out))
    return out, crit


def normed_difference(a, b) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    return lib().oracle_normed_difference(_dp(a), _dp(b), a.size)


def percentage_change(a, b) -> float:
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    return lib().oracle_percentage_change(_ip(a), _ip(b), a.size)


def map_to_grid(grid, centroids):
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    centroids = np.ascontiguousarray(centroids, dtype=np.float64)
    out = np.empty(centroids.shape[0], dtype=np.int32)
    dups = lib().oracle_map_to_grid(_dp(grid), grid.shape[0], _dp(centroids), centroids.shape[0], _ip(out))
    return out, dups


def kmeans(grid, rho, k, max_it, threshold=-1.0, seed=0, init_idx=None, keep_maps=False):
    """KMeansSolver::kernel at one rank.  Returns a dict with interp_idx,
    diffs, pcts, iters, status, final centroids and (optionally) every
    iteration's grid_map."""
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
    n = grid.shape[0]
    idx = np.zeros(k, dtype=np.int32)
    diffs = np.zeros(max(max_it, 1))
    pcts = np.zeros(max(max_it, 1))
    maps = np.zeros((max_it, n), dtype=np.int32) if keep_maps else None
    cen = np.zeros((k, 3))
    iters = ctypes.c_int(0)
    init = None
    if init_idx is not None:
        init = np.ascontiguousarray(init_idx, dtype=np.int32)
    status = lib().oracle_kmeans(_dp(grid), _dp(rho), n, k, max_it, threshold, seed,
                                 _ip(init) if init is not None else None, _ip(idx), _dp(diffs), _dp(pcts),
                                 _ip(maps) if maps is not None else None, _dp(cen), ctypes.byref(iters))
    it = iters.value
    return {"interp_idx": idx, "diffs": diffs[:it], "pcts": pcts[:it], "iters": it,
            "status": status, "centroids": cen, "maps": maps[:it] if maps is not None else None}
