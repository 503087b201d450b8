from ott_amd.initializers.linear import initializers

__all__ = ["initializers"]
