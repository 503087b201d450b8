"""Learned warm starts (the role of ``ott.initializers.neural`` in OTT)."""
