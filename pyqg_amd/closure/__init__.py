"""Learned subgrid closure: a BF16 CNN trained on coarse-grained high-resolution
runs of this solver and used as a q parameterization of the coarse model."""
