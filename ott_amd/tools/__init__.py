from ott_amd.tools import segment_sinkhorn

__all__ = ["segment_sinkhorn"]
