"""Training and evaluation of the DeepSets warm start.

Targets are potentials the Sinkhorn solver itself produces: each training
problem is solved to a tight threshold on the GPU (float32 kernel), and the
network learns the potential of the second cloud, g, up to a per-problem
constant (centred loss); the initializer swaps the clouds to predict f.  Training
uses the project's HIP forward / backward / Adam kernels (``ottx_ds_*``).
The network's output is in units of sigma^2 = mean(C) / 2, so the target of a
problem with epsilon = c mean(C) is g* / sigma^2 = 2 c g* / epsilon.

Problem families (numpy ``default_rng`` seeds):
  gmm      two 4-component Gaussian mixtures in d = 16, n = m = 2048
           (means ~ N(0, 4 I/d), covariance I/d, weights ~ Dirichlet(1));
  gauss    x ~ N(0, s_x^2 I/d), y ~ N(delta, s_y^2 I/d) with a random shift
           delta (|delta| in [0, 0.6]) and scales s in [0.7, 1.4], d = 64.
epsilon = c * mean(C) with c drawn in [0.03, 0.1] (0.05 for evaluation).

  python -m ott_amd.initializers.neural.train --kind gauss --d 64 --n 16384 --count 2000 \
      --epochs 30 --out ott_amd/initializers/neural/weights/deepsets_d64.npz
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from typing import Dict, List, Tuple

import numpy as np
import torch

from ott_amd import datasets
from ott_amd.geometry import pointcloud
from ott_amd.initializers.neural import deepsets
from ott_amd.problems.linear import linear_problem
from ott_amd.solvers.linear import sinkhorn


def make_problem(kind: str, seed: int, n: int, m: int, d: int, eps_mult: float = None
                 ) -> Tuple[np.ndarray, np.ndarray, float]:
  rng = np.random.default_rng(seed)
  if kind == "gmm":
    # the same draws as datasets.gmm_pair(seed, n, m, d)
    x = datasets.gaussian_mixture(rng, n, d, 4)
    y = datasets.gaussian_mixture(rng, m, d, 4)
  elif kind == "gauss":
    sx, sy = rng.uniform(0.7, 1.4, size=2)
    direction = rng.normal(size=d)
    direction /= np.linalg.norm(direction)
    delta = direction * rng.uniform(0.0, 0.6)
    x = rng.normal(size=(n, d)) * sx / np.sqrt(d)
    y = rng.normal(size=(m, d)) * sy / np.sqrt(d) + delta
  else:
    raise ValueError(kind)
  c = float(rng.uniform(0.03, 0.1)) if eps_mult is None else eps_mult
  return x.astype(np.float32), y.astype(np.float32), c


def solve(x, y, eps_mult: float, threshold: float, max_iterations: int = 3000,
          init=None, precision: str = "fp32", inner_iterations: int = 10):
  geom = pointcloud.PointCloud(x, y, epsilon=eps_mult, relative_epsilon="mean", precision=precision)
  prob = linear_problem.LinearProblem(geom)
  out = sinkhorn.Sinkhorn(threshold=threshold, max_iterations=max_iterations,
                          inner_iterations=inner_iterations)(prob, init=init)
  return out, geom.epsilon


def generate(kind: str, seeds: List[int], n: int, m: int, d: int, threshold: float,
             log=print) -> Dict[str, torch.Tensor]:
  """Problems and solver targets (g* / sigma^2), kept on the GPU."""
  dev = torch.device("cuda", torch.cuda.current_device())
  X = torch.empty((len(seeds), n, d), dtype=torch.float32, device=dev)
  Y = torch.empty((len(seeds), m, d), dtype=torch.float32, device=dev)
  T = torch.empty((len(seeds), m), dtype=torch.float32, device=dev)
  E = torch.empty((len(seeds),), dtype=torch.float32)
  R = torch.empty((len(seeds),), dtype=torch.float32)
  iters = []
  t0 = time.time()
  for k, s in enumerate(seeds):
    x, y, c = make_problem(kind, s, n, m, d)
    X[k] = torch.from_numpy(x).to(dev)
    Y[k] = torch.from_numpy(y).to(dev)
    out, eps = solve(X[k], Y[k], c, threshold)
    T[k] = out.g * (2.0 * c / eps)
    E[k] = eps
    R[k] = c
    iters.append(out.n_iters)
    if (k + 1) % max(1, len(seeds) // 10) == 0:
      log(f"  solved {k + 1}/{len(seeds)} problems ({time.time() - t0:.1f} s, "
          f"mean iterations {np.mean(iters):.1f})")
  return dict(X=X, Y=Y, T=T, E=E, R=R, iters=torch.tensor(iters))


def _batch_of(data, idx, max_pool=None) -> Tuple[deepsets.Batch, torch.Tensor]:
  probs = [(data["X"][i], data["Y"][i], float(data["E"][i]), float(data["R"][i])) for i in idx]
  return deepsets.make_batch(probs, max_pool=max_pool), data["T"][idx].reshape(-1)


def train(model: deepsets.DeepSets, data, epochs: int, batch: int, lr: float, seed: int = 0,
          log=print, max_pool=None) -> List[float]:
  rng = np.random.default_rng(seed)
  count = data["X"].shape[0]
  steps_per_epoch = max(1, count // batch)
  total = epochs * steps_per_epoch
  history = []
  step = 0
  t0 = time.time()
  for ep in range(epochs):
    perm = rng.permutation(count)
    losses = []
    for s in range(steps_per_epoch):
      idx = torch.as_tensor(perm[s * batch:(s + 1) * batch], device=data["X"].device)
      bt, target = _batch_of(data, idx.tolist(), max_pool)
      loss, grad = model.loss_backward(bt, target)
      # cosine learning-rate decay
      cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
      model.adam_step(grad, lr=cur)
      losses.append(float(loss.mean()))
      step += 1
    history.append(float(np.mean(losses)))
    log(f"  epoch {ep + 1}/{epochs}: loss {history[-1]:.5f} ({time.time() - t0:.1f} s)")
  return history


def evaluate(model: deepsets.DeepSets, kind: str, seeds: List[int], n: int, m: int, d: int,
             precision: str = "bf16_split", inner_iterations: int = 1, log=print, max_pool=2048) -> Dict:
  """Iterations and time to 1e-3 with zero and learned initialisation."""
  init = deepsets.DeepSetsInitializer(model, max_pool=max_pool)
  rows = []
  for s in seeds:
    x, y, _ = make_problem(kind, s, n, m, d, eps_mult=0.05)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    res = {}
    for name in ("zeros", "learned"):
      torch.cuda.synchronize()
      t0 = time.perf_counter()
      geom = pointcloud.PointCloud(xt, yt, epsilon=0.05, relative_epsilon="mean", precision=precision)
      prob = linear_problem.LinearProblem(geom)
      out = sinkhorn.Sinkhorn(threshold=1e-3, inner_iterations=inner_iterations, max_iterations=5000,
                              initializer=None if name == "zeros" else init)(prob)
      torch.cuda.synchronize()
      res[name] = dict(ms=(time.perf_counter() - t0) * 1e3, n_iters=out.n_iters,
                       reg=out.reg_ot_cost, converged=out.converged)
    rows.append(res)
  summary = {}
  for name in ("zeros", "learned"):
    summary[name] = dict(
        mean_iters=float(np.mean([r[name]["n_iters"] for r in rows])),
        mean_ms=float(np.mean([r[name]["ms"] for r in rows])),
        converged=int(sum(r[name]["converged"] for r in rows)))
  summary["reg_rel_diff_max"] = float(max(abs(r["learned"]["reg"] - r["zeros"]["reg"]) /
                                          abs(r["zeros"]["reg"]) for r in rows))
  summary["problems"] = len(rows)
  log(f"  eval ({len(rows)} problems, {precision}, inner={inner_iterations}): {json.dumps(summary)}")
  return summary


def main(argv=None) -> int:
  ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
  ap.add_argument("--kind", default="gauss", choices=["gauss", "gmm"])
  ap.add_argument("--d", type=int, default=64)
  ap.add_argument("--n", type=int, default=2048)
  ap.add_argument("--count", type=int, default=2000)
  ap.add_argument("--seed0", type=int, default=1000)
  ap.add_argument("--eval-seed0", type=int, default=100000)
  ap.add_argument("--eval-count", type=int, default=32)
  ap.add_argument("--eval-n", type=int, default=0, help="points per cloud at evaluation (0: --n)")
  ap.add_argument("--threshold", type=float, default=1e-5)
  ap.add_argument("--epochs", type=int, default=20)
  ap.add_argument("--batch", type=int, default=16)
  ap.add_argument("--lr", type=float, default=1e-3)
  ap.add_argument("--max-pool", type=int, default=2048, help="encoder points per cloud (0: all)")
  ap.add_argument("--out", required=True)
  args = ap.parse_args(argv)
  log = lambda s: print(s, flush=True)  # noqa: E731
  log(f"generating {args.count} {args.kind} problems (d={args.d}, n=m={args.n})")
  data = generate(args.kind, list(range(args.seed0, args.seed0 + args.count)), args.n, args.n,
                  args.d, args.threshold, log)
  model = deepsets.DeepSets(args.d, seed=0)
  log("training")
  pool = args.max_pool or None
  hist = train(model, data, args.epochs, args.batch, args.lr, log=log, max_pool=pool)
  os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
  model.save(args.out, kind=args.kind, count=args.count, epochs=args.epochs, n=args.n,
             max_pool=args.max_pool, final_loss=hist[-1],
             target_iters_mean=float(data["iters"].float().mean()))
  log(f"saved {args.out}")
  en = args.eval_n or args.n
  seeds = list(range(args.eval_seed0, args.eval_seed0 + args.eval_count))
  # free the training set before evaluating at full size
  del data
  torch.cuda.empty_cache()
  evaluate(model, args.kind, seeds, en, en, args.d, log=log, max_pool=pool)
  return 0


if __name__ == "__main__":
  sys.exit(main())
