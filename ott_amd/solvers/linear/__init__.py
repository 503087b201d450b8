"""Linear OT solvers (``ott.solvers.linear``)."""
from ott_amd.solvers.linear import acceleration, sinkhorn
from ott_amd.solvers.linear._solve import solve

__all__ = ["acceleration", "sinkhorn", "solve"]
