"""Drop-in for the reference's scripts/ops.py: the same op names, from libxdn.so.

    from ops import cross_regression, spatial_filter, spatial_filter_grad, temporal_screening, Reproject
"""
from xdn.ops import (Reproject, cross_regression, spatial_filter, spatial_filter_grad,  # noqa: F401
                     temporal_screening)
