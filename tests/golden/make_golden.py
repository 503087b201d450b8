"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference (JAX) cannot run in this pipeline and its own tests hold no
golden values for this path, so these fixtures are oracle outputs on seeded
inputs (regenerated from ``ott_amd.datasets``): they let GPU tests and the
benchmark check full-size results without re-running a minutes-long FP64 solve.

  python tests/golden/make_golden.py [uniform_d2] [gaussian_d64] [mixture_d128] [batched_d32] [gmm_d16]
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import sinkhorn64 as oracle  # noqa: E402
from ott_amd import datasets  # noqa: E402


def _run(pb, max_iterations, **kw):
  x64 = pb.x.astype(np.float64)
  y64 = pb.y.astype(np.float64)
  n, m = x64.shape[0], y64.shape[0]
  a = np.full(n, 1.0 / n) if pb.a is None else pb.a.astype(np.float64)
  b = np.full(m, 1.0 / m) if pb.b is None else pb.b.astype(np.float64)
  eps = oracle.resolve_epsilon(x64, y64, pb.eps_mult, "mean")
  ops = oracle.PointCloudLSE(x64, y64)
  t0 = time.time()
  out = oracle.sinkhorn(ops, a, b, eps, tau_a=pb.tau_a, tau_b=pb.tau_b,
                        max_iterations=max_iterations, **kw)
  wall = time.time() - t0
  rec = dict(name=pb.name, n=n, m=m, d=int(x64.shape[1]), epsilon=eps,
             tau_a=pb.tau_a, tau_b=pb.tau_b, threshold=out.threshold,
             reg_ot_cost=out.reg_ot_cost, converged=out.converged, n_iters=out.n_iters,
             errors=[float(e) for e in out.errors[:max(1, out.n_iters // 10)]],
             f_mean=float(np.mean(out.f[np.isfinite(out.f)])),
             g_mean=float(np.mean(out.g[np.isfinite(out.g)])),
             oracle_wall_s=wall, oracle_threads=int(oracle.PointCloudLSE(x64[:1], y64[:1])
                                                   ._lib.ox_max_threads()))
  return rec, out


def main(which):
  if "uniform_d2" in which:
    pb = datasets.uniform_shifted(1024, 1024, 2, seed=0)
    rec, out = _run(pb, 2000)
    np.savez_compressed(os.path.join(HERE, "uniform_shifted_n1024_d2.npz"),
                        f=out.f, g=out.g, errors=out.errors,
                        reg_ot_cost=np.array(out.reg_ot_cost), epsilon=np.array(rec["epsilon"]))
    json.dump(rec, open(os.path.join(HERE, "uniform_shifted_n1024_d2.json"), "w"), indent=1)
    print(rec["name"], rec["n_iters"], rec["reg_ot_cost"], flush=True)
  if "gaussian_d64" in which:
    pb = datasets.gaussian_shifted(16384, 16384, 64, seed=1)
    rec, out = _run(pb, 5000)
    json.dump(rec, open(os.path.join(HERE, "gaussian_shifted_n16384_d64.json"), "w"), indent=1)
    print(rec["name"], rec["n_iters"], rec["reg_ot_cost"], rec["oracle_wall_s"], flush=True)
  if "mixture_d128" in which:
    pb = datasets.mixture_pair()
    rec, out = _run(pb, 2000)
    json.dump(rec, open(os.path.join(HERE, "mixture_pair_n32768_m8192_d128.json"), "w"), indent=1)
    print(rec["name"], rec["n_iters"], rec["reg_ot_cost"], rec["oracle_wall_s"], flush=True)
  if "batched_d32" in which:
    # 512 uneven problems; each solved alone, error checked every iteration
    probs = datasets.batched_uneven()
    recs = []
    t0 = time.time()
    for pb in probs:
      rec, _ = _run(pb, 2000, inner_iterations=1)
      recs.append(dict(n=rec["n"], m=rec["m"], epsilon=rec["epsilon"], reg_ot_cost=rec["reg_ot_cost"],
                       n_iters=rec["n_iters"], converged=rec["converged"]))
    json.dump(dict(name="batched_uneven_512_d32", threshold=1e-3, inner_iterations=1, problems=recs,
                   oracle_wall_s=time.time() - t0),
              open(os.path.join(HERE, "batched_uneven_512_d32.json"), "w"))
    print("batched_uneven_512_d32", sum(r["n_iters"] for r in recs), time.time() - t0, flush=True)
  if "gmm_d16" in which:
    # held-out problems of the d=16 warm-start family: zero-start oracle solves,
    # error checked every iteration as in the benchmark
    recs = []
    for seed in range(200000, 200008):
      pb = datasets.gmm_pair(seed)
      rec, _ = _run(pb, 5000, inner_iterations=1)
      recs.append(dict(seed=seed, epsilon=rec["epsilon"], reg_ot_cost=rec["reg_ot_cost"],
                       n_iters=rec["n_iters"], converged=rec["converged"]))
      print(pb.name, rec["n_iters"], rec["reg_ot_cost"], flush=True)
    json.dump(dict(name="gmm_pair_n2048_d16_heldout", threshold=1e-3, inner_iterations=1, problems=recs),
              open(os.path.join(HERE, "gmm_pair_n2048_d16_heldout.json"), "w"), indent=1)


if __name__ == "__main__":
  main(sys.argv[1:] or ["uniform_d2", "gaussian_d64"])
