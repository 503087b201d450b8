from ott_amd.tools import segment_sinkhorn, sinkhorn_divergence

__all__ = ["segment_sinkhorn", "sinkhorn_divergence"]
