"""The tile cache accepts only records that name a variant of the loaded
library (CPU: ottx_tile_configs is a host query)."""

import json
import os

import pytest

from ott_amd import _lib, tune


@pytest.fixture
def cache(tmp_path, monkeypatch):
  if not os.path.exists(_lib.LIB_PATH):
    pytest.skip("libottx.so not built")
  path = tmp_path / "tune.json"
  monkeypatch.setenv("OTTX_TUNE_CACHE", str(path))
  monkeypatch.setattr(tune, "DEFAULTS_PATH", str(tmp_path / "none.json"))
  monkeypatch.setattr(tune, "_cache", None)

  def write(records):
    path.write_text(json.dumps(records))
    tune._cache = None
  return write


def test_valid_record_is_used(cache):
  cfg = tune.tile_configs("bf16_split", 64)[3]
  cache({tune.key("bf16_split", 64, 5000, 5000): tune.entry(cfg)})
  assert tune.lookup("bf16_split", 64, 5000, 5000) == cfg["tile_config"]
  assert tune.lookup("bf16_split", 64, 50000, 5000) == -1  # another size class


def test_stale_records_fall_back_to_the_default(cache):
  table = tune.tile_configs("bf16_split", 64)
  k = tune.key("bf16_split", 64, 5000, 5000)
  # the position now holds another variant
  rec = tune.entry(table[3])
  rec["nb"] = rec["nb"] * 2 + 1
  cache({k: rec})
  assert tune.lookup("bf16_split", 64, 5000, 5000) == -1
  # another ABI version
  rec = dict(tune.entry(table[3]), abi=_lib.ABI_VERSION - 1)
  cache({k: rec})
  assert tune.lookup("bf16_split", 64, 5000, 5000) == -1
  # a position past the table
  rec = dict(tune.entry(table[-1]), tile_config=len(table) + 5)
  cache({k: rec})
  assert tune.lookup("bf16_split", 64, 5000, 5000) == -1
  # a bare index written by an older build
  cache({k: int(table[3]["tile_config"])})
  assert tune.lookup("bf16_split", 64, 5000, 5000) == -1


def test_shipped_defaults_name_listed_variants():
  if not os.path.exists(_lib.LIB_PATH):
    pytest.skip("libottx.so not built")
  with open(tune.DEFAULTS_PATH) as fh:
    shipped = json.load(fh)
  for k, rec in shipped.items():
    _, prec, d = k.split("/")[:3]
    assert tune._valid(rec, prec, int(d[1:])), k


def test_batch_records_are_keyed_by_size_class(cache):
  cfg = tune.tile_configs("bf16_split", 32)[5]
  cache({tune.batch_key("bf16_split", 32, 512, 5.35e8): tune.entry(cfg)})
  assert tune.lookup_batch("bf16_split", 32, 512, 5.35e8) == cfg["tile_config"]
  assert tune.lookup_batch("bf16_split", 32, 500, 5.0e8) == cfg["tile_config"]  # same classes
  assert tune.lookup_batch("bf16_split", 32, 2048, 5.35e8) == -1
  assert tune.lookup_batch("fp32", 32, 512, 5.35e8) == -1
  # single-problem records are not batch records
  assert tune.lookup("bf16_split", 32, 512, 512) == -1
