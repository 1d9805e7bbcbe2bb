"""Learned subgrid closure: a BF16 CNN trained on coarse-grained high-resolution
runs of this solver and used as a q parameterization of the coarse model.

Importing pyqg_amd._lib first keeps a single HIP runtime in the process
whether or not torch is already imported."""
from .. import _lib  # noqa: F401
