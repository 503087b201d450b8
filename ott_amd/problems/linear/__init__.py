from ott_amd.problems.linear import linear_problem, potentials

__all__ = ["linear_problem", "potentials"]
