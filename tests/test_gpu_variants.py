"""Every tile configuration the library lists (``ottx_tile_configs``) against
the FP64 oracle, so that no variant the autotuner can pick -- the 8-wave
ping-pong ones included -- runs without a parity test; and the autotuner
itself returns a listed configuration whose solve matches the oracle.

Each configuration runs three fixed iterations (scan mode, an error every
iteration), which exercises both half-step directions, the fused error pass
and the split-K merge; zero weights exercise the -inf entries.
"""

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "bf16_split": 1e-3}
DIMS = [2, 16, 32, 64, 100, 128]


def _data(d, n=700, m=900):
  rng = np.random.default_rng(d)
  x = (rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32)
  y = (rng.normal(size=(m, d)) / np.sqrt(d) + 0.2 / np.sqrt(d)).astype(np.float32)
  a = rng.uniform(0.2, 1.2, n)
  b = rng.uniform(0.2, 1.2, m)
  a[5] = 0.0
  b[11] = 0.0
  return x, y, (a / a.sum()).astype(np.float32), (b / b.sum()).astype(np.float32)


def _close(got, want, tol, what):
  got = np.asarray(got, np.float64)
  fin = np.isfinite(want)
  assert np.array_equal(fin, np.isfinite(got)), what
  scale = max(1.0, float(np.abs(want[fin]).max()))
  err = float(np.abs(got[fin] - want[fin]).max())
  assert err <= tol * scale, f"{what}: {err:.3e} > {tol:.1e} * {scale:.3e}"


KW = dict(inner_iterations=1, min_iterations=3, max_iterations=3, threshold=1e-3)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("d", DIMS)
def test_every_tile_config_matches_oracle(precision, d, monkeypatch):
  from ott_amd import tune
  from ott_amd.geometry import pointcloud
  from ott_amd.solvers import linear
  x, y, a, b = _data(d)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, relative_epsilon="mean", **KW)
  xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
  at, bt = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
  cfgs = tune.tile_configs(precision, d)
  assert cfgs
  tol = TOL[precision]
  for cfg in cfgs:
    print(f"{precision} d={d} {cfg}", flush=True)
    monkeypatch.setattr(tune, "lookup", lambda *args, c=cfg["tile_config"]: c)
    geom = pointcloud.PointCloud(xt, yt, epsilon=0.05, relative_epsilon="mean", precision=precision)
    out = linear.solve(geom, at, bt, **KW)
    assert out.n_iters == 3
    _close(out.f.cpu().numpy(), ref.f, tol, f"f {cfg}")
    _close(out.g.cpu().numpy(), ref.g, tol, f"g {cfg}")
    np.testing.assert_allclose(out.errors.cpu().numpy(), ref.errors, rtol=20 * tol, atol=1e-6, err_msg=str(cfg))
    assert abs(out.reg_ot_cost - ref.reg_ot_cost) <= tol * abs(ref.reg_ot_cost), cfg


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("d", [16, 64])
def test_every_tile_config_in_the_batched_solve(precision, d):
  """The batched kernel with each variant: three problems of uneven sizes."""
  from ott_amd import tune
  from ott_amd.tools import segment_sinkhorn as ss
  rng = np.random.default_rng(100 + d)
  sizes_x, sizes_y = [300, 517, 64], [411, 129, 700]
  xs = [(rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32) for n in sizes_x]
  ys = [(rng.normal(size=(m, d)) / np.sqrt(d) + 0.1).astype(np.float32) for m in sizes_y]
  refs = [oracle.solve_pointcloud(x, y, epsilon=0.05, relative_epsilon="mean", **KW) for x, y in zip(xs, ys)]
  x = torch.from_numpy(np.concatenate(xs)).cuda()
  y = torch.from_numpy(np.concatenate(ys)).cuda()
  tol = TOL[precision]
  for cfg in tune.tile_configs(precision, d):
    print(f"batched {precision} d={d} {cfg}", flush=True)
    out = ss.segment_sinkhorn_outputs(x, y, num_per_segment_x=sizes_x, num_per_segment_y=sizes_y,
                                      sinkhorn_kwargs=KW, epsilon=0.05, relative_epsilon="mean",
                                      padded_statistics=False, precision=precision,
                                      tile_config=cfg["tile_config"])
    for k, ref in enumerate(refs):
      f = out.f[out.x_offsets[k]:out.x_offsets[k + 1]].cpu().numpy()
      g = out.g[out.y_offsets[k]:out.y_offsets[k + 1]].cpu().numpy()
      _close(f, ref.f, tol, f"f[{k}] {cfg}")
      _close(g, ref.g, tol, f"g[{k}] {cfg}")
      assert abs(float(out.reg_ot_cost[k]) - ref.reg_ot_cost) <= tol * abs(ref.reg_ot_cost), (k, cfg)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_autotune_picks_a_listed_config_that_matches_oracle(precision, tmp_path, monkeypatch):
  from ott_amd import tune
  from ott_amd.geometry import pointcloud
  from ott_amd.solvers import linear
  monkeypatch.setenv("OTTX_TUNE_CACHE", str(tmp_path / "tune.json"))
  monkeypatch.setattr(tune, "_cache", None)
  x, y, a, b = _data(32, n=1500, m=1100)
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=0.05,
                               relative_epsilon="mean", precision=precision)
  res = tune.autotune(geom, reps=2, save=True)
  listed = {c["tile_config"] for c in tune.tile_configs(precision, 32)}
  assert res["best"]["tile_config"] in listed
  assert tune.lookup(precision, 32, 1500, 1100) == res["best"]["tile_config"]
  out = linear.solve(geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), **KW)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, relative_epsilon="mean", **KW)
  _close(out.f.cpu().numpy(), ref.f, TOL[precision], "f")
  assert abs(out.reg_ot_cost - ref.reg_ot_cost) <= TOL[precision] * abs(ref.reg_ot_cost)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_autotune_batch_picks_a_listed_config_that_matches_oracle(precision, tmp_path, monkeypatch):
  """tune.autotune_batch times every variant on the batch itself and records
  the fastest; segment_sinkhorn then runs with it (tile_config -1) and matches
  the oracle."""
  from ott_amd import tune
  from ott_amd.tools import segment_sinkhorn as ss
  monkeypatch.setenv("OTTX_TUNE_CACHE", str(tmp_path / "tune.json"))
  monkeypatch.setattr(tune, "_cache", None)
  sizes = [(300, 257), (64, 900), (1000, 33)]
  rng = np.random.default_rng(9)
  xs = [(rng.normal(size=(n, 16)) / 4).astype(np.float32) for n, _ in sizes]
  ys = [(rng.normal(size=(m, 16)) / 4 + 0.2).astype(np.float32) for _, m in sizes]
  x = torch.from_numpy(np.concatenate(xs)).cuda()
  y = torch.from_numpy(np.concatenate(ys)).cuda()
  nx, ny = [n for n, _ in sizes], [m for _, m in sizes]
  res = tune.autotune_batch(x, y, nx, ny, precision=precision, reps=1)
  listed = {c["tile_config"] for c in tune.tile_configs(precision, 16)}
  assert res["best"]["tile_config"] in listed
  pairs = float(sum(n * m for n, m in sizes))
  assert tune.lookup_batch(precision, 16, 3, pairs) == res["best"]["tile_config"]
  out = ss.segment_sinkhorn_outputs(x, y, num_per_segment_x=nx, num_per_segment_y=ny, sinkhorn_kwargs=KW,
                                    epsilon=0.05, relative_epsilon="mean", padded_statistics=False,
                                    precision=precision)
  tol = TOL[precision]
  for k, (xk, yk) in enumerate(zip(xs, ys)):
    ref = oracle.solve_pointcloud(xk, yk, None, None, epsilon=0.05, relative_epsilon="mean", **KW)
    assert abs(float(out.reg_ot_cost[k]) - ref.reg_ot_cost) <= tol * abs(ref.reg_ot_cost), k
