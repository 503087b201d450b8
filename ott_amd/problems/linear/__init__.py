from ott_amd.problems.linear import linear_problem

__all__ = ["linear_problem"]
