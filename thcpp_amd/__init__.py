"""thcpp_amd: the ISDF factorisation hot path of JoonhoLee-Group/thcpp
(weighted k-means interpolation points, least-squares interpolating vectors,
FFT Coulomb contraction) as hand-written HIP kernels for MI355X (gfx950)
behind the C ABI in include/thcpp_amd.h.

    python -m thcpp_amd.build      # compile libthcpp_amd.so in-tree

thcpp_amd.api.Context wraps the C ABI for Python (torch device tensors);
thcpp_amd.synthetic builds inputs in the orbital-file layout.
"""
from ._lib import LIB_PATH, ThcError, lib  # noqa: F401

__all__ = ["LIB_PATH", "ThcError", "lib"]
