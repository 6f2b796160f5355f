"""Python host over the C ABI, for tests and benchmarks.

Device memory and streams come from PyTorch (plumbing only); every
computation is a call into libthcpp_amd.so.  Method names follow the
reference's functions (src/kmeans.cpp, src/interpolating_vectors.cpp).
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np
import torch

from . import _lib as L

# GEMM tile-configuration table written by the autotuner (python -m thcpp_amd.autotune)
TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "gemm_gfx950.json")
_tuning_applied = [False]


def apply_tuning(path=TUNING_FILE):
    """Install the GEMM tuning table of 'path' (process-wide); a missing file
    clears the table, leaving the built-in rule.  Returns the entry count."""
    entries = []
    if path and os.path.exists(path):
        with open(path) as f:
            entries = json.load(f)["entries"]
    flat = (ctypes.c_int * (5 * len(entries)))(*[int(v) for e in entries for v in e])
    L.check(L.lib().thc_gemm_tuning(flat if entries else None, len(entries)), "thc_gemm_tuning")
    _tuning_applied[0] = True
    return len(entries)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t, dtype, device):
    if isinstance(t, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(t))
    return t.to(device=device, dtype=dtype).contiguous()


class Context:
    """One thc_ctx on one GPU (device index 'device')."""

    def __init__(self, device: int = 0, graphs: bool = False, stream=None):
        """stream: a torch.cuda.Stream or raw hipStream_t handle for the library
        to run on.  By default it is torch's current stream on the device; when
        that is the legacy default stream (handle 0) the library creates a
        blocking stream of its own, which HIP orders after torch's work there.
        The first context of a process installs the committed GEMM tuning
        table (apply_tuning).  graphs: replay the single-system k-means loop
        as a HIP graph (thc_ctx_set_graphs; off by default)."""
        if not _tuning_applied[0]:
            apply_tuning()
        self.device = torch.device("cuda", device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        elif isinstance(stream, torch.cuda.Stream):
            stream = stream.cuda_stream
        self._stream = int(stream or 0)
        h = ctypes.c_void_p()
        L.check(L.lib().thc_ctx_create(device, ctypes.c_void_p(self._stream or None), ctypes.byref(h)),
                "thc_ctx_create")
        self.h = h
        L.check(L.lib().thc_ctx_set_graphs(self.h, 1 if graphs else 0), "thc_ctx_set_graphs")

    def _lib(self):
        """The library, after making torch's pending work on its current stream
        visible to the context's stream: nothing to do when they are the same
        stream or torch's is the legacy default stream (the context's own
        stream is blocking); otherwise wait for torch's stream."""
        cur = torch.cuda.current_stream(self.device).cuda_stream
        if cur != 0 and cur != self._stream:
            torch.cuda.current_stream(self.device).synchronize()
        return L.lib()

    def gemm_time(self, m, n, k, kind=0, cfg=0, reps=3):
        """Device ms of one FP64 GEMM on tile configuration cfg (thc_gemm_time);
        None when cfg does not apply to the shape."""
        ms = ctypes.c_double()
        st = self._lib().thc_gemm_time(self.h, m, n, k, kind, cfg, reps, ctypes.byref(ms))
        if st == L.THC_ERR_UNSUPPORTED:
            return None
        L.check(st, "thc_gemm_time")
        return ms.value

    def close(self):
        if self.h:
            L.lib().thc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def f64(self, a):
        return _dev(a, torch.float64, self.device)

    def i32(self, a):
        return _dev(a, torch.int32, self.device)

    # ------------------------------------------------------------ k-means
    @staticmethod
    def initial_indices(n: int, k: int, seed: int) -> np.ndarray:
        out = np.empty(k, dtype=np.int32)
        L.check(L.lib().thc_kmeans_initial_indices(n, k, seed, out.ctypes.data_as(ctypes.c_void_p)),
                "thc_kmeans_initial_indices")
        return out

    def classify_grid_points(self, grid, centroids):
        g, c = self.f64(grid), self.f64(centroids)
        out = torch.empty(g.shape[0], dtype=torch.int32, device=self.device)
        L.check(self._lib().thc_classify_grid_points(self.h, _p(g), g.shape[0], _p(c), c.shape[0], _p(out)),
                "thc_classify_grid_points")
        return out

    def update_centroids(self, rho, grid, grid_map, k):
        r, g, m = self.f64(rho).reshape(-1), self.f64(grid), self.i32(grid_map)
        out = torch.empty((k, 3), dtype=torch.float64, device=self.device)
        st = self._lib().thc_update_centroids(self.h, _p(r), _p(g), g.shape[0], _p(m), k, _p(out))
        return out, st

    def map_to_grid(self, grid, centroids):
        g, c = self.f64(grid), self.f64(centroids)
        out = torch.empty(c.shape[0], dtype=torch.int32, device=self.device)
        st = self._lib().thc_map_to_grid(self.h, _p(g), g.shape[0], _p(c), c.shape[0], _p(out))
        L.check(st, "thc_map_to_grid", allow=(L.THC_WARN_DUPLICATE_INDICES,))
        return out, st

    def kmeans(self, grid, rho, k, max_it, threshold=-1.0, seed=0, init_idx=None, want_map=False,
               want_centroids=False):
        g, r = self.f64(grid), self.f64(rho).reshape(-1)
        n = g.shape[0]
        idx = torch.empty(k, dtype=torch.int32, device=self.device)
        gm = torch.empty(n, dtype=torch.int32, device=self.device) if want_map else None
        cen = torch.empty((k, 3), dtype=torch.float64, device=self.device) if want_centroids else None
        diffs = np.zeros(max_it)
        pcts = np.zeros(max_it)
        stats = L.KmeansStats()
        stats.diffs = diffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        stats.pcts = pcts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        init = None
        if init_idx is not None:
            init = np.ascontiguousarray(init_idx, dtype=np.int32)
        st = self._lib().thc_kmeans(self.h, _p(g), _p(r), n, k, max_it, threshold, seed,
                                init.ctypes.data_as(ctypes.c_void_p) if init is not None else None,
                                _p(idx), _p(gm), _p(cen), ctypes.byref(stats))
        it = stats.iterations
        return {"status": st, "interp_idx": idx, "grid_map": gm, "centroids": cen, "iters": it,
                "diffs": diffs[:it].copy(), "pcts": pcts[:it].copy(), "converged": bool(stats.converged),
                "duplicates": stats.duplicates, "interp_zeroed": bool(stats.interp_zeroed),
                "elapsed_ms": stats.elapsed_ms}

    def qrcp_points(self, aoR, k, aoR_half=None):
        """First k QRCP pivots of the orbital-product matrix, ascending, and the
        pivot magnitudes |R_jj| in pivot order (src/qrcp.cpp:100-173)."""
        a = self.f64(aoR)
        ah = self.f64(aoR_half) if aoR_half is not None else None
        n, m = a.shape
        idx = torch.empty(k, dtype=torch.int32, device=self.device)
        rd = np.zeros(k)
        st = self._lib().thc_qrcp_points(self.h, _p(a), _p(ah), n, m, ah.shape[1] if ah is not None else 0, k, _p(idx),
                                     rd.ctypes.data_as(ctypes.c_void_p))
        return st, idx, rd

    def qrcp_points_complex(self, aoR, k, aoR_half=None):
        """qrcp_points for complex orbitals ((N, M) complex arrays)."""
        a = self.f64(self._interleave(aoR))
        ah = self.f64(self._interleave(aoR_half)) if aoR_half is not None else None
        n, m = a.shape[0], a.shape[1] // 2
        idx = torch.empty(k, dtype=torch.int32, device=self.device)
        rd = np.zeros(k)
        st = self._lib().thc_qrcp_points_complex(self.h, _p(a), _p(ah), n, m, ah.shape[1] // 2 if ah is not None else 0,
                                             k, _p(idx), rd.ctypes.data_as(ctypes.c_void_p))
        return st, idx, rd

    # ------------------------------------------------------------ ISDF stages
    def czt(self, aoR, idx):
        a, i = self.f64(aoR), self.i32(idx)
        out = torch.empty((i.shape[0], a.shape[0]), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_czt(self.h, _p(a), a.shape[0], a.shape[1], _p(i), i.shape[0], _p(out)), "thc_czt")
        return out

    @staticmethod
    def _interleave(aoR):
        """(N, M) complex -> (N, 2M) float64, (re, im) per orbital."""
        a = np.ascontiguousarray(aoR, dtype=np.complex128)
        return a.view(np.float64).reshape(a.shape[0], 2 * a.shape[1])

    def czt_complex(self, aoR, idx):
        a, i = self.f64(self._interleave(aoR)), self.i32(idx)
        n, m = a.shape[0], a.shape[1] // 2
        out = torch.empty((i.shape[0], n), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_czt_complex(self.h, _p(a), n, m, _p(i), i.shape[0], _p(out)), "thc_czt_complex")
        return out

    def isdf_complex(self, aoR, idx, fft_grid, coulG, precision=L.PREC_FP64, want_luv=True, want_xi=False):
        a, i, cg = self.f64(self._interleave(aoR)), self.i32(idx), self.f64(coulG).reshape(-1)
        n, m = a.shape[0], a.shape[1] // 2
        k = i.shape[0]
        v = torch.empty((k, k), dtype=torch.float64, device=self.device)
        luv = torch.empty((k, k), dtype=torch.float64, device=self.device) if want_luv else None
        xi = torch.empty((k, n), dtype=torch.float64, device=self.device) if want_xi else None
        opts = L.IsdfOpts(precision, 0, 1 if want_luv else 0)
        stats = L.IsdfStats()
        fg = (ctypes.c_int * 3)(*[int(q) for q in fft_grid])
        st = self._lib().thc_isdf_complex(self.h, _p(a), n, m, _p(i), k, fg, _p(cg), ctypes.byref(opts), _p(v), _p(luv),
                                      _p(xi), ctypes.byref(stats))
        return {"status": st, "V": v, "Luv": luv, "xi": xi, "stage_ms": list(stats.stage_ms)}

    def czt_half(self, aoR, aoR_half, idx):
        a, ah, i = self.f64(aoR), self.f64(aoR_half), self.i32(idx)
        out = torch.empty((i.shape[0], a.shape[0]), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_czt_half(self.h, _p(a), _p(ah), a.shape[0], a.shape[1], ah.shape[1], _p(i),
                                     i.shape[0], _p(out)), "thc_czt_half")
        return out

    def cct(self, czt, idx):
        z, i = self.f64(czt), self.i32(idx)
        out = torch.empty((i.shape[0], i.shape[0]), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_cct(self.h, _p(z), z.shape[1], _p(i), i.shape[0], _p(out)), "thc_cct")
        return out

    def cct_rank(self, cct):
        """Numerical rank of CC^T under MatrixOperations::rank's rule
        (src/matrix_operations.h:713-735) and its largest singular value."""
        c = self.f64(cct)
        rank, smax = ctypes.c_int(0), ctypes.c_double(0.0)
        st = self._lib().thc_cct_rank(self.h, _p(c), c.shape[0], ctypes.byref(rank), ctypes.byref(smax))
        L.check(st, "thc_cct_rank", allow=(L.THC_WARN_RANK_DEFICIENT,))
        return rank.value, smax.value, st

    def cholesky(self, a):
        m = self.f64(a).clone()
        info = ctypes.c_int(0)
        L.check(self._lib().thc_cholesky(self.h, _p(m), m.shape[0], ctypes.byref(info)), "thc_cholesky")
        return m, info.value

    def least_squares(self, cct, czt, refine=0):
        c, z = self.f64(cct), self.f64(czt)
        out = torch.empty_like(z)
        info = ctypes.c_int(0)
        st = self._lib().thc_least_squares(self.h, _p(c), _p(z), z.shape[1], z.shape[0], refine, _p(out),
                                       ctypes.byref(info))
        L.check(st, "thc_least_squares")
        return out

    def fft_vectors(self, x, fft_grid, coulG):
        xx, cg = self.f64(x), self.f64(coulG).reshape(-1)
        n = (ctypes.c_int * 3)(*[int(v) for v in fft_grid])
        g2 = 2 * int(fft_grid[0]) * int(fft_grid[1]) * (int(fft_grid[2]) // 2 + 1)
        out = torch.empty((xx.shape[0], g2), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_fft_vectors(self.h, _p(xx), xx.shape[0], n, _p(cg), _p(out)), "thc_fft_vectors")
        return out

    def construct_muv(self, b):
        bb = self.f64(b)
        out = torch.empty((bb.shape[0], bb.shape[0]), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_construct_muv(self.h, _p(bb), bb.shape[0], bb.shape[1], _p(out)), "thc_construct_muv")
        return out

    def isdf(self, aoR, idx, fft_grid, coulG, precision=L.PREC_FP64, refine=0, want_luv=True, want_xi=False,
             check_rank=False):
        a, i, cg = self.f64(aoR), self.i32(idx), self.f64(coulG).reshape(-1)
        n, m = a.shape
        k = i.shape[0]
        v = torch.empty((k, k), dtype=torch.float64, device=self.device)
        luv = torch.empty((k, k), dtype=torch.float64, device=self.device) if want_luv else None
        xi = torch.empty((k, n), dtype=torch.float64, device=self.device) if want_xi else None
        opts = L.IsdfOpts(precision, refine, 1 if want_luv else 0, 1 if check_rank else 0)
        stats = L.IsdfStats()
        fg = (ctypes.c_int * 3)(*[int(q) for q in fft_grid])
        st = self._lib().thc_isdf(self.h, _p(a), n, m, _p(i), k, fg, _p(cg), ctypes.byref(opts), _p(v), _p(luv),
                              _p(xi), ctypes.byref(stats))
        return {"status": st, "V": v, "Luv": luv, "xi": xi, "cct_info": stats.cct_chol_info,
                "luv_info": stats.luv_chol_info, "cct_rank": stats.cct_rank, "cct_smax": stats.cct_smax,
                "stage_ms": list(stats.stage_ms)}

    def isdf_half(self, aoR, aoR_half, idx, fft_grid, coulG, precision=L.PREC_FP64, refine=0, want_xi=False):
        a, ah, i, cg = self.f64(aoR), self.f64(aoR_half), self.i32(idx), self.f64(coulG).reshape(-1)
        n, m = a.shape
        k = i.shape[0]
        muv = torch.empty((k, k), dtype=torch.float64, device=self.device)
        xi = torch.empty((k, n), dtype=torch.float64, device=self.device) if want_xi else None
        opts = L.IsdfOpts(precision, refine, 0)
        stats = L.IsdfStats()
        fg = (ctypes.c_int * 3)(*[int(q) for q in fft_grid])
        st = self._lib().thc_isdf_half(self.h, _p(a), _p(ah), n, m, ah.shape[1], _p(i), k, fg, _p(cg),
                                   ctypes.byref(opts), _p(muv), _p(xi), ctypes.byref(stats))
        return {"status": st, "Muv": muv, "xi": xi, "cct_info": stats.cct_chol_info,
                "stage_ms": list(stats.stage_ms)}

    def czt_half_complex(self, aoR, aoR_half, idx):
        """Half-rotated ZC^T of complex orbitals: (K, N) complex128 torch tensor."""
        a, ah, i = self.f64(self._interleave(aoR)), self.f64(self._interleave(aoR_half)), self.i32(idx)
        n, m, mo, k = a.shape[0], a.shape[1] // 2, ah.shape[1] // 2, i.shape[0]
        out = torch.empty((k, n, 2), dtype=torch.float64, device=self.device)
        L.check(self._lib().thc_czt_half_complex(self.h, _p(a), _p(ah), n, m, mo, _p(i), k, _p(out)),
                "thc_czt_half_complex")
        return torch.view_as_complex(out)

    def isdf_half_complex(self, aoR, aoR_half, idx, fft_grid, coulG, precision=L.PREC_FP64, refine=0,
                          want_xi=False, check_rank=False):
        """Half-rotated ISDF of complex orbitals: Muv (K, K) and, with want_xi,
        X (K, N), complex128 torch tensors."""
        a, ah = self.f64(self._interleave(aoR)), self.f64(self._interleave(aoR_half))
        i, cg = self.i32(idx), self.f64(coulG).reshape(-1)
        n, m, mo, k = a.shape[0], a.shape[1] // 2, ah.shape[1] // 2, i.shape[0]
        muv = torch.empty((k, k, 2), dtype=torch.float64, device=self.device)
        xi = torch.empty((k, n, 2), dtype=torch.float64, device=self.device) if want_xi else None
        opts = L.IsdfOpts(precision, refine, 0, 1 if check_rank else 0)
        stats = L.IsdfStats()
        fg = (ctypes.c_int * 3)(*[int(q) for q in fft_grid])
        st = self._lib().thc_isdf_half_complex(self.h, _p(a), _p(ah), n, m, mo, _p(i), k, fg, _p(cg),
                                           ctypes.byref(opts), _p(muv), _p(xi), ctypes.byref(stats))
        return {"status": st, "Muv": torch.view_as_complex(muv),
                "xi": torch.view_as_complex(xi) if xi is not None else None, "cct_info": stats.cct_chol_info,
                "cct_rank": stats.cct_rank, "stage_ms": list(stats.stage_ms)}

    def factorize(self, grid, rho, aoR, thc_cfac, max_it, fft_grid, coulG, threshold=-1.0, seed=0,
                  precision=L.PREC_FP64, refine=0, want_luv=True):
        g, r, a, cg = self.f64(grid), self.f64(rho).reshape(-1), self.f64(aoR), self.f64(coulG).reshape(-1)
        n, m = a.shape
        k = thc_cfac * m
        idx = torch.empty(k, dtype=torch.int32, device=self.device)
        v = torch.empty((k, k), dtype=torch.float64, device=self.device)
        luv = torch.empty((k, k), dtype=torch.float64, device=self.device) if want_luv else None
        opts = L.IsdfOpts(precision, refine, 1 if want_luv else 0)
        ks, ist = L.KmeansStats(), L.IsdfStats()
        fg = (ctypes.c_int * 3)(*[int(q) for q in fft_grid])
        st = self._lib().thc_factorize(self.h, _p(g), _p(r), _p(a), n, m, thc_cfac, max_it, threshold, seed, fg,
                                   _p(cg), ctypes.byref(opts), _p(idx), _p(v), _p(luv), ctypes.byref(ks),
                                   ctypes.byref(ist))
        return {"status": st, "interp_idx": idx, "V": v, "Luv": luv, "kmeans_ms": ks.elapsed_ms,
                "iters": ks.iterations, "stage_ms": list(ist.stage_ms)}

    def factorize_batched(self, systems, precision=L.PREC_FP64, refine=0, want_luv=True, max_streams=4):
        """systems: sequence of dicts with grid, rho, aoR, coulG, fft_grid, thc_cfac, max_it and optional
        threshold (default -1) and seed (default 0).  Returns (status, list of per-system result dicts)."""
        keep, outs = [], []
        arr = (L.System * len(systems))()
        for i, sd in enumerate(systems):
            g, r, a = self.f64(sd["grid"]), self.f64(sd["rho"]).reshape(-1), self.f64(sd["aoR"])
            cg = self.f64(sd["coulG"]).reshape(-1)
            n, m = a.shape
            k = int(sd["thc_cfac"]) * m
            idx = torch.empty(k, dtype=torch.int32, device=self.device)
            v = torch.empty((k, k), dtype=torch.float64, device=self.device)
            luv = torch.empty((k, k), dtype=torch.float64, device=self.device) if want_luv else None
            keep += [g, r, a, cg]
            q = arr[i]
            q.grid, q.rho, q.aoR, q.coulG = g.data_ptr(), r.data_ptr(), a.data_ptr(), cg.data_ptr()
            q.n, q.m, q.thc_cfac, q.max_it = n, m, int(sd["thc_cfac"]), int(sd["max_it"])
            q.threshold, q.seed = float(sd.get("threshold", -1.0)), int(sd.get("seed", 0))
            for j in range(3):
                q.fft_grid[j] = int(sd["fft_grid"][j])
            q.interp_idx, q.v = idx.data_ptr(), v.data_ptr()
            q.luv = luv.data_ptr() if luv is not None else None
            outs.append({"interp_idx": idx, "V": v, "Luv": luv})
        opts = L.IsdfOpts(precision, refine, 1 if want_luv else 0)
        st = self._lib().thc_factorize_batched(self.h, len(systems), arr, ctypes.byref(opts), max_streams)
        for i, o in enumerate(outs):
            o["status"] = arr[i].status
            o["iters"] = arr[i].kstats.iterations
            o["kmeans_ms"] = arr[i].kstats.elapsed_ms
            o["stage_ms"] = list(arr[i].istats.stage_ms)
        del keep
        return st, outs
