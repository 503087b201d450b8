from ott_amd.solvers import linear

__all__ = ["linear"]
