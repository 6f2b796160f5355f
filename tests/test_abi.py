"""The C-ABI library loads on a CPU-only host and exports every entry point
include/thcpp_amd.h declares; host-only entry points work without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np

from oracle import kmeans_oracle as KO
from thcpp_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "thcpp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # function declarations: "<type> thc_name(" at the start of a line
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(thc_[a-z0-9_]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    lib = L.lib()
    for n in names:
        assert hasattr(lib, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (thc_\w+)", nm))
    assert set(names) <= exported


def test_binding_covers_header():
    assert set(declared_functions()) == set(L.SIGNATURES)


def test_status_strings():
    assert L.status_string(L.THC_OK) == "ok"
    assert "weight" in L.status_string(L.THC_ERR_EMPTY_CLUSTER)
    assert L.lib().thc_version().decode().startswith("thcpp_amd")


def test_initial_indices_host_entry_point():
    """thc_kmeans_initial_indices runs on the host (libstdc++ shuffle)."""
    # both libstdc++ paths: paired draws (N <= 65535) and the one-draw-per-element
    # path restated with a K-entry window (N > 65535), K from 1 to N
    for n, k, seed in [(32768, 128, 0), (262144, 1024, 0), (5000, 17, 99), (65535, 40, 1), (65536, 40, 2),
                       (65536, 65536, 3), (70001, 100, 9), (262144, 2048, 7), (100000, 1, 5),
                       (300000, 299999, 2**32 - 1), (2097152, 8192, 0)]:
        out = np.empty(k, dtype=np.int32)
        st = L.lib().thc_kmeans_initial_indices(n, k, seed, out.ctypes.data_as(ctypes.c_void_p))
        assert st == L.THC_OK
        np.testing.assert_array_equal(out, KO.initial_indices(n, k, seed))
    bad = np.empty(4, dtype=np.int32)
    assert L.lib().thc_kmeans_initial_indices(3, 4, 0, bad.ctypes.data_as(ctypes.c_void_p)) == L.THC_ERR_ARG
