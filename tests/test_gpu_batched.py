"""GPU parity of the batched solver (segment_sinkhorn) against the oracle.

Each problem of a batch must give what the oracle's restatement of OTT's
segment_sinkhorn gives on the padded problem (tools/segment_sinkhorn.py), and
what one ott_amd solve of that problem gives.  Tolerances as in
test_gpu_parity: fp32 rel 1e-4, bf16 split rel 1e-3.
"""

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "bf16_split": 1e-3}
# Padding every problem to 800 rows (OTT's segment layout) shrinks a relative
# epsilon for the small problems; (1, 3) would put it at ~3e-4 of the cost,
# below float32's noise floor for a 1e-4 threshold, so single points are
# covered with an absolute epsilon in test_batched_matches_single_solves.
SIZES = [(37, 50), (200, 130), (513, 64), (64, 700), (129, 129)]


def _data(seed, sizes, d, shift=0.25):
  rng = np.random.default_rng(seed)
  xs = [(rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32) for n, _ in sizes]
  ys = [(rng.normal(size=(m, d)) / np.sqrt(d) + shift).astype(np.float32) for _, m in sizes]
  return np.concatenate(xs), np.concatenate(ys)


def _seg(out, k, side):
  off = out.x_offsets if side == 0 else out.y_offsets
  v = out.f if side == 0 else out.g
  return v[off[k]:off[k + 1]].cpu().numpy().astype(np.float64)


def _close(got, want, tol, what):
  fin = np.isfinite(want)
  assert np.array_equal(fin, np.isfinite(got)), what
  scale = max(1.0, float(np.abs(want[fin]).max()))
  err = float(np.abs(got[fin] - want[fin]).max())
  assert err <= tol * scale, f"{what}: {err:.3e} > {tol:.1e} * {scale:.3e}"


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_batched_matches_oracle_segment_sinkhorn(precision):
  from ott_amd.tools import segment_sinkhorn as ss
  d = 5
  x, y = _data(0, SIZES, d)
  nx, ny = [s[0] for s in SIZES], [s[1] for s in SIZES]
  P = 800
  kw = dict(threshold=1e-4, inner_iterations=1, max_iterations=3000)
  out = ss.segment_sinkhorn_outputs(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(),
                                    num_per_segment_x=nx, num_per_segment_y=ny, max_measure_size=P,
                                    sinkhorn_kwargs=kw, epsilon=0.05, relative_epsilon="mean",
                                    precision=precision)
  refs = oracle.segment_sinkhorn(x, y, nx, ny, max_measure_size=P, sinkhorn_kwargs=kw, epsilon=0.05,
                                 relative_epsilon="mean")
  # the default weights the solve wrote: 1 / n_k per segment (segment.py:100-104), float64 then float32
  np.testing.assert_array_equal(out.a.cpu().numpy(), np.repeat(1.0 / np.array(nx), nx).astype(np.float32))
  np.testing.assert_array_equal(out.b.cpu().numpy(), np.repeat(1.0 / np.array(ny), ny).astype(np.float32))
  tol = TOL[precision]
  for k, ref in enumerate(refs):
    n, m = SIZES[k]
    assert abs(out.epsilon[k].item() - ref.epsilon) <= 1e-6 * ref.epsilon, k
    rel = abs(out.reg_ot_cost[k].item() - ref.reg_ot_cost) / abs(ref.reg_ot_cost)
    assert rel <= tol, (k, out.reg_ot_cost[k].item(), ref.reg_ot_cost)
    assert bool(out.converged[k]) == ref.converged
    # Padding to 800 rows makes epsilon small for the small problems, so some
    # converge slowly; float32 rounding may then move the threshold crossing
    # by a few iterations (1 %).
    assert abs(int(out.n_iters[k]) - ref.n_iters) <= max(1, ref.n_iters // 100), \
        (k, int(out.n_iters[k]), ref.n_iters)
    if int(out.n_iters[k]) == ref.n_iters:
      _close(_seg(out, k, 0), ref.f[:n], tol, f"f[{k}]")
      _close(_seg(out, k, 1), ref.g[:m], tol, f"g[{k}]")


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("tau", [(1.0, 1.0), (0.9, 0.95)])
def test_batched_matches_single_solves(precision, tau):
  from ott_amd.geometry import pointcloud
  from ott_amd.solvers import linear
  from ott_amd.tools import segment_sinkhorn as ss
  d = 32
  sizes = [(64, 4096), (1000, 900), (3000, 64), (257, 511), (1, 3)]
  x, y = _data(1, sizes, d)
  nx, ny = [s[0] for s in sizes], [s[1] for s in sizes]
  kw = dict(threshold=1e-3, inner_iterations=1, max_iterations=2000)
  xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
  out = ss.segment_sinkhorn_outputs(xt, yt, num_per_segment_x=nx, num_per_segment_y=ny,
                                    sinkhorn_kwargs=kw, epsilon=0.1, precision=precision,
                                    tau_a=tau[0], tau_b=tau[1])
  ox, oy = out.x_offsets, out.y_offsets
  for k in range(len(sizes)):
    geom = pointcloud.PointCloud(xt[ox[k]:ox[k + 1]], yt[oy[k]:oy[k + 1]], epsilon=0.1, precision=precision)
    one = linear.solve(geom, tau_a=tau[0], tau_b=tau[1], **kw)
    assert int(out.n_iters[k]) == one.n_iters, k
    assert bool(out.converged[k]) == one.converged
    rel = abs(out.reg_ot_cost[k].item() - one.reg_ot_cost) / abs(one.reg_ot_cost)
    assert rel <= 1e-5, (k, rel)
    np.testing.assert_allclose(out.errors[k].cpu().numpy(), one.errors.cpu().numpy(), rtol=1e-3, atol=1e-7)
    _close(_seg(out, k, 0), one.f.cpu().numpy().astype(np.float64), 1e-5, f"f[{k}]")


def test_batched_segment_ids_and_weights():
  from ott_amd.tools import segment_sinkhorn as ss
  rng = np.random.default_rng(3)
  d = 3
  sizes = [(40, 30), (10, 60), (25, 25)]
  x, y = _data(4, sizes, d)
  nx, ny = [s[0] for s in sizes], [s[1] for s in sizes]
  wx = rng.uniform(0.1, 1.0, size=x.shape[0]).astype(np.float32)
  wy = rng.uniform(0.1, 1.0, size=y.shape[0]).astype(np.float32)
  wx[3] = 0.0  # a zero weight gives a -inf potential
  ox, oy = np.cumsum([0] + nx), np.cumsum([0] + ny)
  for k in range(3):
    wx[ox[k]:ox[k + 1]] /= wx[ox[k]:ox[k + 1]].sum()
    wy[oy[k]:oy[k + 1]] /= wy[oy[k]:oy[k + 1]].sum()
  kw = dict(threshold=1e-4, inner_iterations=1)
  base = ss.segment_sinkhorn(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), num_per_segment_x=nx,
                             num_per_segment_y=ny, weights_x=torch.from_numpy(wx).cuda(),
                             weights_y=torch.from_numpy(wy).cuda(), sinkhorn_kwargs=kw, epsilon=0.1)
  ids_x, ids_y = np.repeat(np.arange(3), nx), np.repeat(np.arange(3), ny)
  px, py = rng.permutation(x.shape[0]), rng.permutation(y.shape[0])
  shuf = ss.segment_sinkhorn(torch.from_numpy(x[px]).cuda(), torch.from_numpy(y[py]).cuda(), num_segments=3,
                             max_measure_size=100, segment_ids_x=torch.from_numpy(ids_x[px]),
                             segment_ids_y=torch.from_numpy(ids_y[py]),
                             weights_x=torch.from_numpy(wx[px]).cuda(), weights_y=torch.from_numpy(wy[py]).cuda(),
                             sinkhorn_kwargs=kw, epsilon=0.1)
  # the order within a segment changes the reduction order only
  torch.testing.assert_close(shuf, base, rtol=1e-5, atol=0)
  for k in range(3):
    sl = slice(ox[k], ox[k + 1])
    ref = oracle.solve_pointcloud(x[sl], y[oy[k]:oy[k + 1]], wx[sl], wy[oy[k]:oy[k + 1]], epsilon=0.1, **kw)
    assert abs(base[k].item() - ref.reg_ot_cost) <= 1e-4 * abs(ref.reg_ot_cost)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_batched_thousands_of_problems_retire_in_rounds(precision):
  """9000 problems (three templates repeated, each with its own epsilon, so
  they stop at different blocks): the live work lists are rebuilt on the
  device as problems finish, over more items than one 8192-item round of
  the rebuild.  Every copy must give its template's result bit for bit, and
  the template the oracle's (tools/segment_sinkhorn.py:24-138 per problem)."""
  from ott_amd.tools import segment_sinkhorn as ss
  d = 4
  templates = [(20, 30, 0.5), (45, 17, 0.1), (33, 60, 0.05)]
  reps = 3000
  rng = np.random.default_rng(11)
  tx = [(rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32) for n, _, _ in templates]
  ty = [(rng.normal(size=(m, d)) / np.sqrt(d) + 0.3).astype(np.float32) for _, m, _ in templates]
  order = np.tile(np.arange(3), reps)
  x = np.concatenate([tx[t] for t in order])
  y = np.concatenate([ty[t] for t in order])
  nx = [templates[t][0] for t in order]
  ny = [templates[t][1] for t in order]
  eps = [templates[t][2] for t in order]
  kw = dict(threshold=1e-4, inner_iterations=1, max_iterations=3000)
  print(f"batched {len(order)} problems {precision}", flush=True)
  out = ss.segment_sinkhorn_outputs(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), num_per_segment_x=nx,
                                    num_per_segment_y=ny, sinkhorn_kwargs=kw, epsilons=eps, precision=precision)
  f = out.f.cpu().numpy()
  g = out.g.cpu().numpy()
  reg = out.reg_ot_cost.numpy()
  iters = out.n_iters.numpy()
  ox, oy = out.x_offsets, out.y_offsets
  tol = TOL[precision]
  for t, (n, m, e) in enumerate(templates):
    ks = np.nonzero(order == t)[0]
    k0 = ks[0]
    for k in ks[1:]:
      assert reg[k] == reg[k0] and iters[k] == iters[k0], (t, k)
      assert np.array_equal(f[ox[k]:ox[k + 1]], f[ox[k0]:ox[k0 + 1]], equal_nan=True), (t, k)
      assert np.array_equal(g[oy[k]:oy[k + 1]], g[oy[k0]:oy[k0 + 1]], equal_nan=True), (t, k)
    ref = oracle.solve_pointcloud(tx[t], ty[t], None, None, epsilon=e, **kw)
    assert ref.converged and bool(out.converged[k0])
    assert iters[k0] == ref.n_iters, (t, iters[k0], ref.n_iters)
    assert abs(reg[k0] - ref.reg_ot_cost) <= tol * abs(ref.reg_ot_cost), (t, reg[k0], ref.reg_ot_cost)
    _close(f[ox[k0]:ox[k0 + 1]].astype(np.float64), ref.f, tol, f"f[{t}]")
  # the templates stop at different blocks, so the lists shrink in steps
  assert len(set(int(v) for v in iters[:3])) == 3, iters[:3]
