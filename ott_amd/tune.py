"""Tile autotuner for the fused half-step kernel.

OTT leaves tiling to XLA; here the kernel exposes a set of tile
configurations (``ottx_tile_configs``: 16-column blocks per wave, 16-row
blocks per streamed tile, waves per workgroup, target workgroups per CU).
``autotune`` times every configuration with ``ottx_benchmark_halfstep`` on the
problem at hand (both half-step directions) and records the fastest in a JSON
cache keyed by device, precision, feature dimension and the power-of-two size
class of n and m.  ``lookup`` returns the cached choice, falling back to the
defaults shipped in ``tuned_configs.json``, then to -1 (the kernel default).

A tile_config is a position in the library's variant table, so every entry
also records the variant it names ({tile_config, nb, mt, w, target}) and the
ABI version; ``lookup`` accepts an entry only if the loaded library still lists
exactly that variant at that position.  A cache written by another build
(whose table differs) therefore falls back to the default instead of sending
solves to a variant the entry did not measure.

The cache path is ``$OTTX_TUNE_CACHE`` or ``~/.cache/ott_amd/tune.json``.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import threading
from typing import Dict, List, Optional

import torch

from ott_amd import _lib

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS_PATH = os.path.join(_HERE, "tuned_configs.json")
_lock = threading.Lock()
_cache: Optional[Dict[str, dict]] = None


def cache_path() -> str:
  return os.environ.get("OTTX_TUNE_CACHE",
                        os.path.join(os.path.expanduser("~"), ".cache", "ott_amd", "tune.json"))


def _device_name() -> str:
  try:
    return torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]
  except Exception:  # noqa: BLE001 - name only keys the cache
    return "gfx950"


def _size_class(v: int) -> int:
  return int(math.ceil(math.log2(max(int(v), 1))))


def key(precision: str, d: int, n: int, m: int) -> str:
  return f"{_device_name()}/{precision}/d{d}/n{_size_class(n)}/m{_size_class(m)}"


def tile_configs(precision: str, d: int) -> List[dict]:
  lib = _lib.load_library()
  prec = _lib.PRECISIONS[precision]
  count = lib.ottx_tile_configs(prec, d, None, 0)
  buf = (ctypes.c_int32 * (5 * count))()
  lib.ottx_tile_configs(prec, d, buf, count)
  return [dict(tile_config=buf[5 * i], nb=buf[5 * i + 1], mt=buf[5 * i + 2], w=buf[5 * i + 3],
               target=buf[5 * i + 4]) for i in range(count)]


_FIELDS = ("tile_config", "nb", "mt", "w", "target")


def entry(cfg: dict) -> dict:
  """The cache record of a variant (as ``tile_configs`` lists it)."""
  return dict({f: int(cfg[f]) for f in _FIELDS}, abi=_lib.ABI_VERSION)


def _load() -> Dict[str, dict]:
  global _cache
  if _cache is None:
    merged: Dict[str, dict] = {}
    for path in (DEFAULTS_PATH, cache_path()):
      if os.path.exists(path):
        try:
          with open(path) as fh:
            merged.update({k: v for k, v in json.load(fh).items() if isinstance(v, dict)})
        except (OSError, ValueError):
          pass
    _cache = merged
    _resolved.clear()  # validated lookups of the previous cache
  return _cache


def _valid(rec: dict, precision: str, d: int) -> bool:
  """The record names a variant this library lists at the same position."""
  try:
    if int(rec.get("abi", -1)) != _lib.ABI_VERSION:
      return False
    want = {f: int(rec[f]) for f in _FIELDS}
  except (KeyError, TypeError, ValueError):
    return False
  return any(all(c[f] == want[f] for f in _FIELDS) for c in tile_configs(precision, d))


_resolved: Dict[str, int] = {}


def _resolve(k: str, precision: str, d: int) -> int:
  """The validated tile_config of cache key ``k`` (-1 when none), memoised:
  the loaded library's variant table does not change within a process, and a
  solve on the hot path should not pay for the check every call."""
  with _lock:
    rec = _load().get(k)  # reloads (and forgets the memo) after a reset
    hit = _resolved.get(k)
  if hit is not None:
    return hit
  val = int(rec["tile_config"]) if rec is not None and _valid(rec, precision, d) else -1
  with _lock:
    _resolved[k] = val
  return val


def lookup(precision: str, d: int, n: int, m: int) -> int:
  """The tuned tile_config for the shape, or -1 when none is recorded or the
  record does not name a variant of the loaded library."""
  return _resolve(key(precision, d, n, m), precision, d)


def batch_key(precision: str, d: int, count: int, pairs: float) -> str:
  """Cache key of a batched solve: precision, d, and the power-of-two classes
  of the problem count and of the total pairs per half-step."""
  return f"{_device_name()}/{precision}/d{d}/batch/k{_size_class(count)}/p{_size_class(int(pairs))}"


def lookup_batch(precision: str, d: int, count: int, pairs: float) -> int:
  """The tuned tile_config of a batched solve, or -1 (the variant with the
  least row padding) when none is recorded or it names no listed variant."""
  return _resolve(batch_key(precision, d, count, pairs), precision, d)


def _record(k: str, best: dict, save: bool) -> None:
  with _lock:
    _load()[k] = entry(best)
    _resolved.pop(k, None)
    if save:
      path = cache_path()
      os.makedirs(os.path.dirname(path), exist_ok=True)
      existing = {}
      if os.path.exists(path):
        try:
          with open(path) as fh:
            existing = json.load(fh)
        except (OSError, ValueError):
          existing = {}
      existing[k] = entry(best)
      with open(path, "w") as fh:
        json.dump(existing, fh, indent=1, sort_keys=True)


def autotune_batch(x, y, num_per_segment_x, num_per_segment_y, precision: str = "bf16_split", reps: int = 3,
                   iterations: int = 2, save: bool = True, verbose: bool = False, **pointcloud_kwargs) -> dict:
  """Time every tile configuration on a batched solve (segment_sinkhorn of the
  given segments, ``iterations`` fixed iterations so that convergence does
  not enter) and record the fastest under :func:`batch_key`."""
  import time
  from ott_amd.tools import segment_sinkhorn as ss
  d = int(x.shape[1])
  kw = dict(min_iterations=iterations, max_iterations=iterations, inner_iterations=1)
  pointcloud_kwargs.setdefault("epsilon", 0.05)
  pointcloud_kwargs.setdefault("relative_epsilon", "mean")
  results = []
  for cfg in tile_configs(precision, d):
    def one():
      return ss.segment_sinkhorn_outputs(x, y, num_per_segment_x=num_per_segment_x,
                                         num_per_segment_y=num_per_segment_y, sinkhorn_kwargs=kw,
                                         precision=precision, tile_config=cfg["tile_config"], **pointcloud_kwargs)
    one()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
      torch.cuda.synchronize()
      t0 = time.perf_counter()
      one()
      torch.cuda.synchronize()
      times.append((time.perf_counter() - t0) * 1e3)
    results.append(dict(cfg, ms=min(times)))
    if verbose:
      print(f"  batch tile {cfg}: {min(times):.3f} ms", flush=True)
  best = min(results, key=lambda r: r["ms"])
  pairs = float(sum(int(n) * int(m) for n, m in zip(num_per_segment_x, num_per_segment_y)))
  k = batch_key(precision, d, len(num_per_segment_x), pairs)
  _record(k, best, save)
  return dict(best=best, key=k, results=results)


def benchmark(geom, tile_config: int, axis: int, reps: int = 10) -> float:
  """Mean milliseconds of one half-step of ``geom`` with ``tile_config``."""
  lib = _lib.load_library()
  prob, keep = geom._problem()
  # the workspace of the configuration being timed
  opts = _lib.SinkhornOpts(threshold=1e-3, norm_error=1, inner_iterations=10, min_iterations=0,
                           max_iterations=10, momentum=1.0, tile_config=tile_config, use_graph=0)
  nbytes = lib.ottx_sinkhorn_workspace_bytes(ctypes.byref(prob), ctypes.byref(opts))
  ws = _lib.workspace(nbytes, geom.device)
  n, m = geom.shape
  f = torch.zeros(n, device=geom.device)
  g = torch.zeros(m, device=geom.device)
  avg = ctypes.c_double(0.0)
  _lib.check("ottx_benchmark_halfstep", lib.ottx_benchmark_halfstep(
      ctypes.byref(prob), axis, tile_config, reps, f.data_ptr(), g.data_ptr(),
      ctypes.byref(avg), ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
  del keep
  return float(avg.value)


def autotune(geom, reps: int = 10, save: bool = True, verbose: bool = False) -> dict:
  """Time every tile configuration on ``geom``; record and return the fastest."""
  n, m = geom.shape
  d = geom.x.shape[1]
  results = []
  for cfg in tile_configs(geom.precision, d):
    ms = benchmark(geom, cfg["tile_config"], 0, reps) + benchmark(geom, cfg["tile_config"], 1, reps)
    results.append(dict(cfg, ms_per_iteration=ms))
    if verbose:
      print(f"  tile {cfg}: {ms:.4f} ms per iteration", flush=True)
  best = min(results, key=lambda r: r["ms_per_iteration"])
  k = key(geom.precision, d, n, m)
  _record(k, best, save)
  return dict(best=best, key=k, results=results)


def main(argv=None) -> int:
  """Autotune a list of shapes: ``python -m ott_amd.tune 64x16384x16384 2x1024x1024 ...``.

  Each shape is ``d x n x m``; uniform-cube points (seeded) are used since the
  kernel's time does not depend on the values.  Results go to the cache file
  (``$OTTX_TUNE_CACHE``) and are printed as JSON.
  """
  import argparse
  import numpy as np
  from ott_amd.geometry import pointcloud
  ap = argparse.ArgumentParser(description="tile autotuner")
  ap.add_argument("shapes", nargs="+")
  ap.add_argument("--precisions", default="bf16_split,fp32")
  ap.add_argument("--reps", type=int, default=10)
  args = ap.parse_args(argv)
  out = {}
  for shape in args.shapes:
    d, n, m = (int(v) for v in shape.split("x"))
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.uniform(size=(n, d)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.uniform(size=(m, d)).astype(np.float32)).cuda()
    for prec in args.precisions.split(","):
      geom = pointcloud.PointCloud(x, y, epsilon=0.05, relative_epsilon="mean", precision=prec)
      res = autotune(geom, reps=args.reps)
      worst = max(r["ms_per_iteration"] for r in res["results"])
      default = [r for r in res["results"] if r["tile_config"] == res["results"][0]["tile_config"]][0]
      out[res["key"]] = dict(best=res["best"], default_ms=default["ms_per_iteration"], worst_ms=worst)
      print(json.dumps({res["key"]: out[res["key"]]}), flush=True)
  return 0


if __name__ == "__main__":
  raise SystemExit(main())
