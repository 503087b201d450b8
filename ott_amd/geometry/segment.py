"""Segmentation of concatenated point clouds (geometry/segment.py:23-190).

``segment_point_cloud`` reproduces OTT's padded layout ([num_segments,
max_measure_size, d] with zero-weight padding rows) for callers that want it.
The batched solver itself does not pad: ``contiguous_segments`` groups the
rows of each segment together and returns offsets, which is what
``ottx_sinkhorn_batched_solve`` takes.
"""

from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np
import torch

__all__ = ["segment_point_cloud", "contiguous_segments"]


def _counts_and_order(num: int, num_segments: Optional[int], segment_ids, indices_are_sorted: bool,
                      num_per_segment: Optional[Sequence[int]]):
  """Per-segment sizes and the row order that makes segments contiguous."""
  if segment_ids is not None:
    if num_segments is None:
      raise ValueError("Please specify `num_segments`.")
    ids = torch.as_tensor(segment_ids).to("cpu", torch.int64)
    if ids.shape != (num,):
      raise ValueError(f"segment_ids must have shape [{num}], got {tuple(ids.shape)}")
    if ids.numel() and (int(ids.min()) < 0 or int(ids.max()) >= num_segments):
      raise ValueError("segment ids must lie in range(num_segments)")
    counts = torch.bincount(ids, minlength=num_segments).numpy()
    # rows of segment i in their original order (segment.py:106-108)
    order = None if indices_are_sorted else torch.argsort(ids, stable=True)
    return counts.astype(np.int64), order
  if num_per_segment is None:
    raise ValueError("Please specify `num_per_segment`.")
  counts = np.asarray(list(num_per_segment), dtype=np.int64)
  if num_segments is not None and num_segments != len(counts):
    raise ValueError("num_segments does not match len(num_per_segment)")
  if int(counts.sum()) != num:
    raise ValueError(f"num_per_segment sums to {int(counts.sum())}, not to {num} rows")
  return counts, None


def contiguous_segments(x: torch.Tensor, a: Optional[torch.Tensor] = None, num_segments: Optional[int] = None,
                        segment_ids=None, indices_are_sorted: bool = False,
                        num_per_segment: Optional[Sequence[int]] = None, default_weights: bool = True
                        ) -> Tuple[torch.Tensor, Optional[torch.Tensor], np.ndarray]:
  """Rows grouped by segment, their weights and the offsets [K + 1].

  Weights default to ``1 / n_k`` within each segment (segment.py:100-104);
  with ``default_weights=False`` absent weights stay ``None`` (the batched
  solve then writes them on the device).
  """
  num = x.shape[0]
  counts, order = _counts_and_order(num, num_segments, segment_ids, indices_are_sorted, num_per_segment)
  if np.any(counts < 1):
    raise ValueError("every segment must hold at least one point")
  offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
  if order is not None:
    idx = order.to(x.device)
    x = x.index_select(0, idx)
    if a is not None:
      a = a.index_select(0, idx)
  if a is None and default_weights:
    per = torch.from_numpy(np.repeat(1.0 / counts, counts).astype(np.float32))
    a = per.to(x.device)
  return x.contiguous(), (None if a is None else a.contiguous()), offsets


def segment_point_cloud(x, a=None, num_segments: Optional[int] = None, max_measure_size: Optional[int] = None,
                        segment_ids=None, indices_are_sorted: bool = False,
                        num_per_segment: Optional[Sequence[int]] = None, padding_vector=None
                        ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
  """geometry/segment.py:23-124: padded [K, P, d] points, [K, P] weights, sizes [K].

  As in OTT, ``max_measure_size`` is capped at the number of rows, and rows
  beyond it in a segment are dropped; padding rows carry zero weight.
  """
  x = torch.as_tensor(x)
  num, dim = x.shape
  if segment_ids is not None and max_measure_size is None:
    raise ValueError("Please specify `max_measure_size`.")
  counts, order = _counts_and_order(num, num_segments, segment_ids, indices_are_sorted, num_per_segment)
  K = len(counts)
  if max_measure_size is None:
    max_measure_size = int(counts.max())
  P = min(int(max_measure_size), num)
  if a is None:
    a = torch.from_numpy(np.repeat(1.0 / np.maximum(counts, 1), counts).astype(np.float32)).to(x.device)
  a = torch.as_tensor(a, device=x.device)
  if order is not None:
    x = x.index_select(0, order.to(x.device))
    a = a.index_select(0, order.to(x.device))
  pad = torch.zeros((1, dim), dtype=x.dtype, device=x.device) if padding_vector is None else \
      torch.as_tensor(padding_vector, dtype=x.dtype, device=x.device).reshape(1, dim)
  seg_x = pad.expand(K, P, dim).clone()
  seg_a = torch.zeros((K, P), dtype=a.dtype, device=x.device)
  off = 0
  for k in range(K):
    c = int(counts[k])
    take = min(c, P)
    seg_x[k, :take] = x[off:off + take]
    seg_a[k, :take] = a[off:off + take]
    off += c
  return seg_x, seg_a, torch.as_tensor(counts)
