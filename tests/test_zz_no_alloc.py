"""ottx_pointcloud_op_noalloc (static device scratch, no allocation) against
ottx_pointcloud_op (caller workspace) through the C ABI: identical outputs on
two calls in a row, the second on a smaller problem, and a clear error for a
problem beyond the scratch."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu


def _problem(lib_mod, x, y, precision, eps):
  n, d = x.shape
  m = y.shape[0]
  return lib_mod.Problem(n=n, m=m, d=d, precision=lib_mod.PRECISIONS[precision], x=x.data_ptr(), y=y.data_ptr(),
                         a=None, b=None, epsilon=eps, inv_scale_cost=1.0, tau_a=1.0, tau_b=1.0,
                         relative_epsilon=0.0)


def _both(n, m, d, precision, seed):
  from ott_amd import _lib
  lib = _lib.load_library()
  rng = np.random.default_rng(seed)
  xh = (rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32)
  yh = (rng.normal(size=(m, d)) / np.sqrt(d) + 0.1).astype(np.float32)
  fh = (rng.normal(size=n) * 0.05).astype(np.float32)
  gh = (rng.normal(size=m) * 0.05).astype(np.float32)
  x, y = torch.from_numpy(xh).cuda(), torch.from_numpy(yh).cuda()
  f, g = torch.from_numpy(fh).cuda(), torch.from_numpy(gh).cuda()
  eps = 0.07
  pb = _problem(_lib, x, y, precision, eps)
  need = lib.ottx_pointcloud_op_noalloc_bytes(ctypes.byref(pb))
  assert 0 < need <= _lib.NOALLOC_SCRATCH_BYTES
  ws = torch.empty(lib.ottx_op_workspace_bytes(ctypes.byref(pb)), dtype=torch.uint8, device="cuda")
  stream = _lib.stream_ptr()
  results = []
  for op, axis in ((_lib.OP_APPLY_LSE, 0), (_lib.OP_APPLY_LSE, 1), (_lib.OP_UPDATE, 1), (_lib.OP_MARGINAL, 0)):
    n_out = m if axis == 0 else n
    lm = torch.full((n_out,), -2.0, device="cuda") if op == _lib.OP_UPDATE else None
    lm_ptr = None if lm is None else lm.data_ptr()
    want = torch.empty(n_out, device="cuda")
    got = torch.empty(n_out, device="cuda")
    print(f"noalloc n={n} m={m} d={d} {precision} op={op} axis={axis}", flush=True)
    assert lib.ottx_pointcloud_op(ctypes.byref(pb), op, axis, f.data_ptr(), g.data_ptr(), lm_ptr, eps,
                                  want.data_ptr(), ws.data_ptr(), ws.numel(), stream) == 0
    assert lib.ottx_pointcloud_op_noalloc(ctypes.byref(pb), op, axis, f.data_ptr(), g.data_ptr(), lm_ptr, eps,
                                          got.data_ptr(), stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(got, want), (op, axis, float((got - want).abs().max()))
    results.append((op, axis, got.cpu().numpy()))
  # one operator checked against the oracle too
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(xh, yh))
  ref = oracle.apply_lse_kernel(ops, fh.astype(np.float64), gh.astype(np.float64), eps, 0)
  tol = {"fp32": 1e-4, "bf16_split": 1e-3}[precision]
  assert np.abs(results[0][2] - ref).max() <= tol * max(1.0, np.abs(ref).max())


def test_noalloc_matches_workspace_entry_twice_second_smaller():
  _both(3000, 2500, 64, "bf16_split", seed=1)
  _both(700, 500, 16, "fp32", seed=2)


def test_noalloc_rejects_problems_beyond_the_scratch():
  from ott_amd import _lib
  lib = _lib.load_library()
  pb = _lib.Problem(n=400000, m=400000, d=64, precision=1, x=256, y=256, epsilon=0.1, inv_scale_cost=1.0,
                    tau_a=1.0, tau_b=1.0, relative_epsilon=0.0)
  assert lib.ottx_pointcloud_op_noalloc_bytes(ctypes.byref(pb)) > _lib.NOALLOC_SCRATCH_BYTES
  # rejected before any launch (the pointers are never dereferenced)
  rc = lib.ottx_pointcloud_op_noalloc(ctypes.byref(pb), _lib.OP_APPLY_LSE, 0, 256, 256, None, 0.1, 256,
                                      _lib.stream_ptr())
  assert rc == _lib.OTTX_ESCRATCH
  assert b"scratch" in lib.ottx_strerror(rc)
