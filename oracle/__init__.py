"""CPU oracle for the log-domain PointCloud Sinkhorn path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports it.  Parity notes for each function live in the module docstrings.
"""
