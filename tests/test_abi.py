"""The C-ABI library loads and exports every symbol declared in include/*.h.

Runs without a GPU: only symbol lookup and pure host queries, no kernel launch.
"""

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
  names = set()
  for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
    text = open(h).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names |= set(re.findall(r"\b(ottx_\w+)\s*\(", text))
  return names


def test_header_declares_entry_points():
  names = _declared()
  for required in ("ottx_sinkhorn_solve", "ottx_pointcloud_op", "ottx_mean_cost"):
    assert required in names


def test_library_exports_every_declared_symbol():
  from ott_amd import _lib
  if not os.path.exists(_lib.LIB_PATH):
    pytest.skip("libottx.so not built")
  lib = _lib.load_library()
  missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
  assert not missing, f"missing exports: {missing}"
  assert set(_lib.EXPORTED_SYMBOLS) == _declared()
  assert lib.ottx_abi_version() == 5
  assert lib.ottx_strerror(-1) == b"invalid argument"


def test_struct_layouts_match_header():
  from ott_amd import _lib
  assert ctypes.sizeof(_lib.Problem) == 8 + 8 + 4 + 4 + 4 * 8 + 5 * 8 + 2 * 4 + 8
  assert ctypes.sizeof(_lib.SinkhornOpts) == 88 + 2 * 8
  assert ctypes.sizeof(_lib.SinkhornOut) == 3 * 8 + 3 * 8 + 4 * 4
  assert ctypes.sizeof(_lib.Batch) == 4 * 4 + 4 * 8 + 3 * 8 + 6 * 8
  assert ctypes.sizeof(_lib.BatchOut) == 3 * 8 + 4 * 8 + 2 * 4
  assert ctypes.sizeof(_lib.DsBatch) == 4 * 4 + 10 * 8 + 2 * 4


def test_batched_workspace_query_rejects_bad_batches():
  from ott_amd import _lib
  lib = _lib.load_library()
  x_off = (ctypes.c_int64 * 3)(0, 5, 12)
  y_off = (ctypes.c_int64 * 3)(0, 7, 9)
  eps = (ctypes.c_double * 2)(0.1, 0.2)
  dummy = 256
  def batch(**kw):
    args = dict(count=2, d=3, precision=0, x=dummy, y=dummy, a=dummy, b=dummy, x_off=x_off, y_off=y_off,
                epsilon=eps, relative_epsilon=0.0, mean_rows_x=0.0, mean_rows_y=0.0, inv_scale_cost=1.0,
                tau_a=1.0, tau_b=1.0)
    args.update(kw)
    return _lib.Batch(**args)
  opts = _lib.SinkhornOpts(threshold=1e-3, norm_error=1, inner_iterations=10, min_iterations=0,
                           max_iterations=100, momentum=1.0, tile_config=-1, use_graph=1)
  assert lib.ottx_sinkhorn_batched_workspace_bytes(ctypes.byref(batch()), ctypes.byref(opts)) > 0
  bad_off = (ctypes.c_int64 * 3)(0, 5, 5)  # an empty problem
  # the default-weight flags are accepted, unknown flag bits are not
  for flags in (_lib.BATCH_UNIFORM_A, _lib.BATCH_UNIFORM_B, _lib.BATCH_UNIFORM_A | _lib.BATCH_UNIFORM_B):
    assert lib.ottx_sinkhorn_batched_workspace_bytes(ctypes.byref(batch(flags=flags)), ctypes.byref(opts)) > 0
  for kw in (dict(count=0), dict(d=0), dict(x_off=bad_off), dict(epsilon=None), dict(mean_rows_x=4.0),
             dict(tau_a=0.0), dict(precision=7), dict(flags=4), dict(flags=-1)):
    assert lib.ottx_sinkhorn_batched_workspace_bytes(ctypes.byref(batch(**kw)), ctypes.byref(opts)) == 0, kw
  # relative epsilon needs no per-problem epsilons
  assert lib.ottx_sinkhorn_batched_workspace_bytes(
      ctypes.byref(batch(epsilon=None, relative_epsilon=0.05, mean_rows_x=7.0, mean_rows_y=7.0)),
      ctypes.byref(opts)) > 0


def test_invalid_arguments_are_rejected_without_gpu_work():
  from ott_amd import _lib
  if not os.path.exists(_lib.LIB_PATH):
    pytest.skip("libottx.so not built")
  lib = _lib.load_library()
  pb = _lib.Problem(n=0, m=4, d=2)
  assert lib.ottx_sinkhorn_workspace_bytes(ctypes.byref(pb), None) == 0
  assert lib.ottx_op_workspace_bytes(ctypes.byref(pb)) == 0
  assert lib.ottx_mean_cost_workspace_bytes(0, 3, 2) == 0
  assert lib.ottx_cost_moments_workspace_bytes(0, 3, 2) == 0
  assert lib.ottx_cost_moments_workspace_bytes(3, 3, 257) == 0
  assert lib.ottx_cost_moments_workspace_bytes(3, 3, 2) > 0
  # an unknown epsilon statistic is rejected at planning
  bad = _lib.Problem(n=8, m=8, d=2, precision=0, x=1, y=1, epsilon=0.1, inv_scale_cost=1.0, tau_a=1.0,
                     tau_b=1.0, relative_epsilon=0.05, epsilon_statistic=2)
  assert lib.ottx_op_workspace_bytes(ctypes.byref(bad)) == 0
  rc = lib.ottx_sinkhorn_solve(ctypes.byref(pb), None, None, None, 0, None)
  assert rc == _lib.OTTX_EINVAL


@pytest.mark.parametrize("n,m,d", [(257, 129, 32), (1, 1, 1), (17, 29, 4), (5000, 3, 256)])
@pytest.mark.parametrize("precision", [0, 1])
def test_workspace_planning_all_shapes(n, m, d, precision):
  """Host-side planning of every tile configuration (no kernel launch)."""
  from ott_amd import _lib
  if not os.path.exists(_lib.LIB_PATH):
    pytest.skip("libottx.so not built")
  lib = _lib.load_library()
  pb = _lib.Problem(n=n, m=m, d=d, precision=precision, x=1, y=1, a=1, b=1, epsilon=0.1,
                    inv_scale_cost=1.0, tau_a=1.0, tau_b=1.0, relative_epsilon=0.0)
  opts = _lib.SinkhornOpts(threshold=1e-3, norm_error=1, inner_iterations=10, min_iterations=0,
                           max_iterations=100, momentum=1.0, tile_config=-1, use_graph=1)
  assert lib.ottx_op_workspace_bytes(ctypes.byref(pb)) > 0
  assert lib.ottx_sinkhorn_workspace_bytes(ctypes.byref(pb), ctypes.byref(opts)) > 0
  assert lib.ottx_tile_configs(precision, d, None, 0) > 0
  # a relative epsilon on std(C) adds the moment partials to the workspace
  std = _lib.Problem(n=n, m=m, d=d, precision=precision, x=1, y=1, a=1, b=1, epsilon=0.0,
                     inv_scale_cost=1.0, tau_a=1.0, tau_b=1.0, relative_epsilon=0.05,
                     epsilon_statistic=_lib.EPS_STAT_STD)
  assert (lib.ottx_sinkhorn_workspace_bytes(ctypes.byref(std), ctypes.byref(opts)) >
          lib.ottx_sinkhorn_workspace_bytes(ctypes.byref(pb), ctypes.byref(opts)))
