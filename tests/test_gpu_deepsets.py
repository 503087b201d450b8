"""DeepSets warm-start kernels against a float64 torch-CPU autograd reference.

The HIP forward rounds inputs, weight matrices and hidden activations to
bfloat16 (float32 accumulation); the reference applies the same roundings in
float64, so forward values agree to float32-accumulation level.  Gradients are
compared with float64 autograd of that same function: the kernels additionally
round back-propagated activations to bfloat16, so the bound is a relative L2
error of 3e-2 per tensor (and cosine > 0.999).
"""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf(t):
  return t.to(torch.bfloat16).to(torch.float64)


def _reference(params, layout, problems, target, bf16=True, pool=None):
  """Float64 forward + loss of the network; returns (g list, loss, named params)."""
  from ott_amd.initializers.neural import deepsets
  P = {}
  for name, shape in layout.shapes:
    off = dict(layout.offsets)[name]
    n = int(np.prod(shape))
    P[name] = torch.tensor(params[off:off + n], dtype=torch.float64).view(*shape).requires_grad_(True)
  rnd = _bf if bf16 else (lambda t: t)
  W = {k: (rnd(P[k].detach()) + (P[k] - P[k].detach())) for k in ("W0", "W1", "W2", "Wy", "W4")}
  outs, losses = [], []
  q = 0
  for prob in problems:
    x, y, eps = prob[:3]
    rel = prob[3] if len(prob) > 3 else None
    x = torch.tensor(x, dtype=torch.float64)
    y = torch.tensor(y, dtype=torch.float64)
    d = x.shape[1]
    # x x^T etc. in float64; the kernel's float32 statistics agree to ~1e-6
    mx, my = x.mean(0), y.mean(0)
    mu = 0.5 * (mx + my)
    meanc = (x * x).sum(1).mean() + (y * y).sum(1).mean() - 2 * (mx @ my)
    sig2 = 0.5 * meanc
    # exact moments of the normalised clouds over every point
    s_x = ((x - mu) ** 2).sum(1).mean() / sig2
    s_y = ((y - mu) ** 2).sum(1).mean() / sig2
    m_y = ((y - mu) / torch.sqrt(sig2)).mean(0)
    def norm(p):
      z = torch.zeros(p.shape[0], layout.kin, dtype=torch.float64)
      pt = ((p - mu) / torch.sqrt(sig2)).to(torch.float32).to(torch.float64)
      z[:, :d] = pt
      z[:, d] = (pt * pt).sum(1) - 1.0
      return rnd(z)
    xt, yt = norm(x), norm(y)
    def enc(p):
      a0 = torch.relu(p @ W["W0"].T + P["b0"])
      a0 = rnd(a0.detach()) + (a0 - a0.detach())
      a1 = torch.relu(a0 @ W["W1"].T + P["b1"])
      a1 = rnd(a1.detach()) + (a1 - a1.detach())
      return a1 @ W["W2"].T + P["b2"]
    def sub(p):  # evenly spaced encoder rows when a cloud exceeds the pool size
      if pool is None or p.shape[0] <= pool:
        return p
      return p[(torch.arange(pool) * p.shape[0]) // pool]
    ex, ey = enc(sub(xt)).mean(0), enc(sub(yt)).mean(0)
    kappa = torch.tensor(2.0 * rel, dtype=torch.float64) if rel is not None else eps / sig2
    width = deepsets.ctx_width(d)
    tail = torch.cat([kappa.reshape(1), (s_x - 1.0).reshape(1), (s_y - 1.0).reshape(1), m_y])
    ctx = torch.cat([ex, ey, tail, torch.zeros(width - 512 - tail.shape[0], dtype=torch.float64)])
    hctx = P["Wc"] @ ctx + P["bh"]
    zq = P["Wq"] @ ctx
    z1 = torch.relu(yt @ W["Wy"].T + hctx)
    z1 = rnd(z1.detach()) + (z1 - z1.detach())
    a2 = torch.relu(z1 @ W["W4"].T + P["b4"])
    out = a2 @ P["w5"] + P["b5"][0] + yt @ zq
    m = y.shape[0]
    r = out - torch.tensor(target[q:q + m], dtype=torch.float64)
    losses.append(((r - r.mean()) ** 2).mean())
    outs.append(sig2 * out)
    q += m
  loss = torch.stack(losses).mean()
  return outs, loss, losses, P


def _problems(d, seed):
  rng = np.random.default_rng(seed)
  shapes = [(50, 70), (130, 40), (64, 64)]
  out = []
  for n, m in shapes:
    x = rng.normal(size=(n, d)).astype(np.float32) / np.sqrt(d)
    y = (rng.normal(size=(m, d)) / np.sqrt(d) + 0.2).astype(np.float32)
    out.append((x, y, float(rng.uniform(0.03, 0.2))))
  return out


@pytest.mark.parametrize("d,relative", [(5, False), (40, False), (63, True)])
def test_forward_and_gradients_match_float64_autograd(d, relative):
  from ott_amd.initializers.neural import deepsets
  model = deepsets.DeepSets(d, seed=3)
  # give the output layer non-zero weights so every gradient path is exercised
  rng = np.random.default_rng(4)
  with torch.no_grad():
    model.tensor("w5").copy_(torch.from_numpy(rng.normal(size=256).astype(np.float32) * 0.05))
    model.tensor("b5")[0] = 0.3
    wq = model.tensor("Wq")
    wq.copy_(torch.from_numpy(rng.normal(size=tuple(wq.shape)).astype(np.float32) * 0.02))
  model.refresh()
  probs = _problems(d, seed=d)
  if relative:
    probs = [(x, y, 0.0, e) for x, y, e in probs]  # only eps / mean cost given
  batch = deepsets.make_batch([(torch.from_numpy(p[0]), torch.from_numpy(p[1]), *p[2:]) for p in probs])
  m_total = sum(p[1].shape[0] for p in probs)
  target = rng.normal(size=m_total).astype(np.float32)

  g = model.forward(batch).cpu().numpy().astype(np.float64)
  params = model.params.cpu().numpy()
  outs, loss_ref, losses_ref, P = _reference(params, model.layout, probs, target)
  g_ref = torch.cat(outs).detach().numpy()
  scale = np.max(np.abs(g_ref))
  assert np.max(np.abs(g - g_ref)) < 2e-3 * scale, np.max(np.abs(g - g_ref)) / scale

  loss, grad = model.loss_backward(batch, torch.from_numpy(target))
  np.testing.assert_allclose(loss.cpu().numpy(), [float(l.detach()) for l in losses_ref], rtol=5e-3)
  loss_ref.backward()
  grad = grad.cpu().numpy().astype(np.float64)
  for name, shape in model.layout.shapes:
    off = dict(model.layout.offsets)[name]
    n = int(np.prod(shape))
    got = grad[off:off + n]
    want = P[name].grad.detach().numpy().reshape(-1)
    if name == "b5":
      # the centred loss is invariant to a constant output shift: zero gradient
      assert abs(got[0]) < 1e-6 * np.linalg.norm(grad) and np.all(got[1:] == 0.0), got[:2]
      continue
    nrm = np.linalg.norm(want)
    if nrm == 0.0:
      assert np.linalg.norm(got) < 1e-6, name
      continue
    rel = np.linalg.norm(got - want) / nrm
    cos = got @ want / (np.linalg.norm(got) * nrm)
    # b4's gradient is w5[k] sum_q dout[q] (a2[q, k] > 0) and the centred loss
    # makes sum_q dout[q] = 0 per problem: a small difference of large sums,
    # where float32 forward values (dout, the ReLU mask) move it by a few percent
    bound = 6e-2 if name == "b4" else 3e-2
    assert rel < bound and cos > 0.999, (name, rel, cos)


def test_adam_step_and_save_load(tmp_path):
  from ott_amd.initializers.neural import deepsets
  d = 8
  model = deepsets.DeepSets(d, seed=1)
  probs = _problems(d, seed=11)
  batch = deepsets.make_batch([(torch.from_numpy(x), torch.from_numpy(y), e) for x, y, e in probs])
  m_total = sum(y.shape[0] for _, y, _ in probs)
  target = torch.from_numpy(np.random.default_rng(2).normal(size=m_total).astype(np.float32))
  p0 = model.params.clone()
  _, grad = model.loss_backward(batch, target)
  model.adam_step(grad, lr=1e-3)
  # first Adam step: p1 = p0 - lr * g / (|g| + eps') with bias correction
  g = grad.double()
  m = 0.1 * g / (1 - 0.9)
  v = 0.001 * g * g / (1 - 0.999)
  want = p0.double() - 1e-3 * m / (torch.sqrt(v) + 1e-8)
  torch.testing.assert_close(model.params.double(), want, rtol=1e-5, atol=1e-7)
  # steps 2 and 3 with unrelated gradients, where the bias corrections
  # 1 - 0.9^t and 1 - 0.999^t no longer cancel (optax.adam: moments
  # m <- 0.9 m + 0.1 g, v <- 0.999 v + 0.001 g^2, update m_hat / (sqrt(v_hat) + eps))
  rng = np.random.default_rng(5)
  m_s, v_s, p = 0.1 * g, 0.001 * g * g, want
  for t in (2, 3):
    gt = torch.from_numpy(rng.normal(scale=float(g.abs().mean()) * 3, size=g.numel())).to(grad.dtype).to(grad.device)
    model.adam_step(gt, lr=1e-3)
    gd = gt.double()
    m_s = 0.9 * m_s + 0.1 * gd
    v_s = 0.999 * v_s + 0.001 * gd * gd
    p = p - 1e-3 * (m_s / (1 - 0.9 ** t)) / (torch.sqrt(v_s / (1 - 0.999 ** t)) + 1e-8)
    torch.testing.assert_close(model.params.double(), p, rtol=1e-5, atol=1e-7)
  path = str(tmp_path / "m.npz")
  model.save(path)
  other = deepsets.DeepSets.load(path)
  torch.testing.assert_close(other.forward(batch), model.forward(batch))


def test_training_reduces_loss():
  """A few Adam steps on a fixed batch reduce the centred MSE."""
  from ott_amd.initializers.neural import deepsets
  d = 4
  model = deepsets.DeepSets(d, seed=0)
  probs = _problems(d, seed=5)
  batch = deepsets.make_batch([(torch.from_numpy(x), torch.from_numpy(y), e) for x, y, e in probs])
  ys = np.concatenate([y for _, y, _ in probs])
  target = torch.from_numpy((ys @ np.arange(1, d + 1) + (ys ** 2).sum(1)).astype(np.float32))
  first = None
  for _ in range(60):
    loss, grad = model.loss_backward(batch, target)
    first = float(loss.mean()) if first is None else first
    model.adam_step(grad, lr=3e-3)
  last = float(model.loss_backward(batch, target)[0].mean())
  assert last < 0.2 * first, (first, last)


@pytest.mark.parametrize("predict", ["g", "f"])
@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_initializer_fast_path_matches_explicit_init(precision, predict):
  """Sinkhorn(initializer=DeepSetsInitializer) hands the prediction to the solve
  (predict="g": f0 derived inside the solve, OTTX_INIT_F_FROM_G; predict="f":
  g set inside the solve, OTTX_INIT_F_GIVEN); it must agree with passing
  init=initializer(prob) computed outside."""
  from ott_amd.geometry import pointcloud
  from ott_amd.initializers.neural import deepsets
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  d = 16
  model = deepsets.DeepSets(d, seed=7)
  rng = np.random.default_rng(8)
  with torch.no_grad():
    model.tensor("w5").copy_(torch.from_numpy(rng.normal(size=256).astype(np.float32) * 0.05))
  model.refresh()
  init = deepsets.DeepSetsInitializer(model, predict=predict)
  x = torch.from_numpy(rng.normal(size=(700, d)).astype(np.float32) / 4).cuda()
  y = torch.from_numpy((rng.normal(size=(513, d)) / 4 + 0.1).astype(np.float32)).cuda()
  outs = []
  for fast in (True, False):
    geom = pointcloud.PointCloud(x, y, epsilon=0.05, relative_epsilon="mean", precision=precision)
    prob = linear_problem.LinearProblem(geom)
    solver = sinkhorn.Sinkhorn(threshold=1e-4, inner_iterations=1, initializer=init)
    outs.append(solver(prob) if fast else solver(prob, init=init(prob)))
  a, b = outs
  assert a.n_iters == b.n_iters
  assert abs(a.reg_ot_cost - b.reg_ot_cost) <= 1e-6 * abs(b.reg_ot_cost)
  torch.testing.assert_close(a.f, b.f, rtol=1e-4, atol=1e-4 * float(b.f.abs().max()))


def test_relative_epsilon_and_absolute_give_same_prediction():
  """kappa from eps / mean cost (unresolved relative epsilon) equals eps / sigma^2."""
  from ott_amd.initializers.neural import deepsets
  d = 8
  model = deepsets.DeepSets(d, seed=2)
  rng = np.random.default_rng(3)
  with torch.no_grad():
    model.tensor("w5").copy_(torch.from_numpy(rng.normal(size=256).astype(np.float32) * 0.05))
    wc = model.tensor("Wc")
    wc.copy_(torch.from_numpy(rng.normal(scale=0.1, size=tuple(wc.shape)).astype(np.float32)))
  model.refresh()
  x = rng.normal(size=(300, d)).astype(np.float32)
  y = (rng.normal(size=(200, d)) + 0.5).astype(np.float32)
  meanc = float(((x[:, None, :].astype(np.float64) - y[None]) ** 2).sum(-1).mean())
  c = 0.07
  g_abs = model.forward(deepsets.make_batch([(torch.from_numpy(x), torch.from_numpy(y), c * meanc)]))
  g_rel = model.forward(deepsets.make_batch([(torch.from_numpy(x), torch.from_numpy(y), 0.0, c)]))
  # the two kappas differ in the last float bits (host float64 mean cost
  # against the device's), which can move a bf16 rounding of a network input
  # by one step: 1e-3 of the largest value, far below what a wrong kappa gives
  torch.testing.assert_close(g_abs, g_rel, rtol=1e-3, atol=1e-3 * float(g_rel.abs().max()))


def test_pooled_subsample_matches_reference():
  """max_pool caps the encoder rows per cloud (evenly spaced); forward and
  gradients follow the same float64 function with the subsample."""
  from ott_amd.initializers.neural import deepsets
  d = 6
  model = deepsets.DeepSets(d, seed=5)
  rng = np.random.default_rng(6)
  with torch.no_grad():
    model.tensor("w5").copy_(torch.from_numpy(rng.normal(size=256).astype(np.float32) * 0.05))
  model.refresh()
  probs = _problems(d, seed=9)  # clouds of 40 to 130 points
  pool = 48
  batch = deepsets.make_batch([(torch.from_numpy(x), torch.from_numpy(y), e) for x, y, e in probs], max_pool=pool)
  assert batch.enc_rows == sum(min(x.shape[0], pool) + min(y.shape[0], pool) for x, y, _ in probs)
  m_total = sum(y.shape[0] for _, y, _ in probs)
  target = rng.normal(size=m_total).astype(np.float32)
  g = model.forward(batch).cpu().numpy().astype(np.float64)
  outs, loss_ref, _, P = _reference(model.params.cpu().numpy(), model.layout, probs, target, pool=pool)
  g_ref = torch.cat(outs).detach().numpy()
  assert np.max(np.abs(g - g_ref)) < 2e-3 * np.max(np.abs(g_ref))
  _, grad = model.loss_backward(batch, torch.from_numpy(target))
  loss_ref.backward()
  grad = grad.cpu().numpy().astype(np.float64)
  for name in ("W0", "W2", "Wc", "w5"):
    off, shape = model.layout.slice(name)
    got = grad[off:off + int(np.prod(shape))]
    want = P[name].grad.detach().numpy().reshape(-1)
    assert np.linalg.norm(got - want) < 3e-2 * np.linalg.norm(want), name
