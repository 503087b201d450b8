from ott_amd.initializers import linear

__all__ = ["linear"]
