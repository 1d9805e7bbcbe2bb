"""Learned subgrid closure: a BF16 CNN trained on coarse-grained high-resolution
runs of this solver and used as a q parameterization of the coarse model.

libqgk is loaded before torch: both bring a libamdhip64.so.7, and the process
must use the ROCm runtime libqgk and libqgcnn were built against."""
from .. import _lib  # noqa: F401  (must precede any torch import)
