from ott_amd.problems import linear

__all__ = ["linear"]
