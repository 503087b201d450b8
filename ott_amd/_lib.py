"""ctypes binding of ``libottx.so`` (the C ABI in ``include/ottx.h``).

The library is built in-tree (``ott_amd/lib/libottx.so``) by ``ott_amd._build``.
There is no fallback: every entry point here raises when the library or a GPU
is missing.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# OTTX_LIBRARY overrides the in-tree build (for A/B comparisons of builds).
LIB_PATH = os.environ.get("OTTX_LIBRARY", os.path.join(_HERE, "lib", "libottx.so"))

OTTX_OK = 0
OTTX_EINVAL = -1
OTTX_EUNSUPPORTED = -2
OTTX_EHIP = -3
OTTX_EWORKSPACE = -4
OTTX_ESCRATCH = -5
APPLY_PREPARED = 1
APPLY_LSE_READY = 2
NOALLOC_SCRATCH_BYTES = 32 << 20  # OTTX_NOALLOC_SCRATCH_BYTES

PREC_FP32 = 0
PREC_BF16_SPLIT = 1
PRECISIONS = {"fp32": PREC_FP32, "bf16_split": PREC_BF16_SPLIT}
INIT_GIVEN, INIT_F_FROM_G, INIT_ZEROS, INIT_F_GIVEN = 0, 1, 2, 3

OP_APPLY_LSE = 0
OP_UPDATE = 1
OP_MARGINAL = 2


class Problem(ctypes.Structure):
  _fields_ = [
      ("n", ctypes.c_int64),
      ("m", ctypes.c_int64),
      ("d", ctypes.c_int32),
      ("precision", ctypes.c_int32),
      ("x", ctypes.c_void_p),
      ("y", ctypes.c_void_p),
      ("a", ctypes.c_void_p),
      ("b", ctypes.c_void_p),
      ("epsilon", ctypes.c_double),
      ("inv_scale_cost", ctypes.c_double),
      ("tau_a", ctypes.c_double),
      ("tau_b", ctypes.c_double),
      ("relative_epsilon", ctypes.c_double),
      ("epsilon_statistic", ctypes.c_int32),
      ("cost", ctypes.c_int32),
      ("cost_p", ctypes.c_double),
  ]


# OTTX_ABI_VERSION of include/ottx.h that these structures mirror
ABI_VERSION = 5

# ottx_batch.flags
BATCH_UNIFORM_A = 1
BATCH_UNIFORM_B = 2

EPS_STAT_MEAN = 0
EPS_STAT_STD = 1

# void (*)(void* user, int32_t block, double error)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double)


class SinkhornOpts(ctypes.Structure):
  _fields_ = [
      ("threshold", ctypes.c_double),
      ("norm_error", ctypes.c_int32),
      ("inner_iterations", ctypes.c_int32),
      ("min_iterations", ctypes.c_int32),
      ("max_iterations", ctypes.c_int32),
      ("momentum", ctypes.c_double),
      ("tile_config", ctypes.c_int32),
      ("use_graph", ctypes.c_int32),
      ("init_mode", ctypes.c_int32),
      ("recenter_potentials", ctypes.c_int32),
      ("parallel_dual_updates", ctypes.c_int32),
      ("epsilon_init", ctypes.c_double),
      ("epsilon_decay", ctypes.c_double),
      ("momentum_start", ctypes.c_int32),
      ("_pad0", ctypes.c_int32),
      ("momentum_error_threshold", ctypes.c_double),
      ("progress_fn", PROGRESS_FN),
      ("progress_user", ctypes.c_void_p),
  ]


class SinkhornOut(ctypes.Structure):
  _fields_ = [
      ("f", ctypes.c_void_p),
      ("g", ctypes.c_void_p),
      ("errors", ctypes.c_void_p),
      ("reg_ot_cost", ctypes.c_double),
      ("epsilon", ctypes.c_double),
      ("total_mass", ctypes.c_double),
      ("converged", ctypes.c_int32),
      ("n_iters", ctypes.c_int32),
      ("n_blocks", ctypes.c_int32),
      ("_pad", ctypes.c_int32),
  ]


class Batch(ctypes.Structure):
  _fields_ = [
      ("count", ctypes.c_int32),
      ("d", ctypes.c_int32),
      ("precision", ctypes.c_int32),
      ("flags", ctypes.c_int32),
      ("x", ctypes.c_void_p),
      ("y", ctypes.c_void_p),
      ("a", ctypes.c_void_p),
      ("b", ctypes.c_void_p),
      ("x_off", ctypes.POINTER(ctypes.c_int64)),
      ("y_off", ctypes.POINTER(ctypes.c_int64)),
      ("epsilon", ctypes.POINTER(ctypes.c_double)),
      ("relative_epsilon", ctypes.c_double),
      ("mean_rows_x", ctypes.c_double),
      ("mean_rows_y", ctypes.c_double),
      ("inv_scale_cost", ctypes.c_double),
      ("tau_a", ctypes.c_double),
      ("tau_b", ctypes.c_double),
  ]


class BatchOut(ctypes.Structure):
  _fields_ = [
      ("f", ctypes.c_void_p),
      ("g", ctypes.c_void_p),
      ("errors", ctypes.c_void_p),
      ("reg_ot_cost", ctypes.POINTER(ctypes.c_double)),
      ("epsilon", ctypes.POINTER(ctypes.c_double)),
      ("converged", ctypes.POINTER(ctypes.c_int32)),
      ("n_iters", ctypes.POINTER(ctypes.c_int32)),
      ("n_blocks", ctypes.c_int32),
      ("_pad", ctypes.c_int32),
  ]


class DsModel(ctypes.Structure):
  _fields_ = [
      ("d", ctypes.c_int32),
      ("_pad", ctypes.c_int32),
      ("params", ctypes.c_void_p),
      ("weights_bf16", ctypes.c_void_p),
  ]


class DsBatch(ctypes.Structure):
  _fields_ = [
      ("count", ctypes.c_int32),
      ("d", ctypes.c_int32),
      ("n_total", ctypes.c_int32),
      ("m_total", ctypes.c_int32),
      ("x", ctypes.c_void_p),
      ("y", ctypes.c_void_p),
      ("eps", ctypes.c_void_p),
      ("x_off", ctypes.c_void_p),
      ("y_off", ctypes.c_void_p),
      ("y_prob", ctypes.c_void_p),
      ("enc_src", ctypes.c_void_p),
      ("seg_of_row", ctypes.c_void_p),
      ("seg_off", ctypes.c_void_p),
      ("eps_rel", ctypes.c_void_p),
      ("enc_rows", ctypes.c_int32),
      ("_pad1", ctypes.c_int32),
  ]


# (name, restype, argtypes) of every exported symbol declared in include/*.h.
_SIGNATURES = [
    ("ottx_abi_version", ctypes.c_int, []),
    ("ottx_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("ottx_mean_cost_workspace_bytes", ctypes.c_size_t,
     [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32]),
    ("ottx_mean_cost", ctypes.c_int,
     [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_cost_moments_workspace_bytes", ctypes.c_size_t,
     [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32]),
    ("ottx_cost_moments", ctypes.c_int,
     [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_op_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(Problem)]),
    ("ottx_pointcloud_op", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_pointcloud_op_noalloc", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    ("ottx_pointcloud_op_noalloc_bytes", ctypes.c_size_t, [ctypes.POINTER(Problem)]),
    ("ottx_apply_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(Problem), ctypes.c_int32]),
    ("ottx_pointcloud_apply", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
      ctypes.c_double, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_cost_maxima", ctypes.c_int,
     [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
      ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_cost_median_workspace_bytes", ctypes.c_size_t, []),
    ("ottx_cost_median", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_cost_summary_workspace_bytes", ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    ("ottx_cost_summary", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_transport_matrix", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("ottx_transport_cost_workspace_bytes", ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    ("ottx_transport_cost", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_size_t,
      ctypes.c_void_p]),
    ("ottx_cost_gradient_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(Problem), ctypes.c_int32]),
    ("ottx_cost_gradient", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_sinkhorn_workspace_bytes", ctypes.c_size_t,
     [ctypes.POINTER(Problem), ctypes.POINTER(SinkhornOpts)]),
    ("ottx_sinkhorn_batched_workspace_bytes", ctypes.c_size_t,
     [ctypes.POINTER(Batch), ctypes.POINTER(SinkhornOpts)]),
    ("ottx_sinkhorn_batched_solve", ctypes.c_int,
     [ctypes.POINTER(Batch), ctypes.POINTER(SinkhornOpts), ctypes.POINTER(BatchOut), ctypes.c_void_p,
      ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_sinkhorn_solve", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.POINTER(SinkhornOpts),
      ctypes.POINTER(SinkhornOut), ctypes.c_void_p, ctypes.c_size_t,
      ctypes.c_void_p]),
    ("ottx_tile_configs", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    ("ottx_benchmark_halfstep", ctypes.c_int,
     [ctypes.POINTER(Problem), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_ds_param_count", ctypes.c_int64, [ctypes.c_int32]),
    ("ottx_ds_weight_bf16_count", ctypes.c_int64, [ctypes.c_int32]),
    ("ottx_ds_workspace_bytes", ctypes.c_size_t, [ctypes.POINTER(DsBatch), ctypes.c_int32]),
    ("ottx_ds_refresh_bf16", ctypes.c_int, [ctypes.POINTER(DsModel), ctypes.c_void_p]),
    ("ottx_ds_forward", ctypes.c_int,
     [ctypes.POINTER(DsModel), ctypes.POINTER(DsBatch), ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_ds_loss_backward", ctypes.c_int,
     [ctypes.POINTER(DsModel), ctypes.POINTER(DsBatch), ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ("ottx_ds_adam_step", ctypes.c_int,
     [ctypes.POINTER(DsModel), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]),
]

EXPORTED_SYMBOLS = tuple(name for name, _, _ in _SIGNATURES)

_lib_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
  """Load ``libottx.so`` and set every prototype; raise if it is not built."""
  global _lib
  with _lib_lock:
    if _lib is not None:
      return _lib
    if not os.path.exists(path):
      raise RuntimeError(
          f"ottx HIP library not found at {path}; build it with "
          "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, restype, argtypes in _SIGNATURES:
      fn = getattr(lib, name)
      fn.restype = restype
      fn.argtypes = argtypes
    if lib.ottx_abi_version() != ABI_VERSION:
      raise RuntimeError(f"{path} has ABI version {lib.ottx_abi_version()}, these bindings expect "
                         f"{ABI_VERSION}; rebuild it")
    _lib = lib
    return lib


class OttxError(RuntimeError):
  """A non-zero status returned by the C ABI."""

  def __init__(self, fn: str, status: int):
    msg = load_library().ottx_strerror(status).decode()
    super().__init__(f"{fn} failed: {msg} ({status})")
    self.status = status


def check(fn: str, status: int) -> None:
  if status == OTTX_EUNSUPPORTED:
    raise NotImplementedError(f"{fn}: option not supported by the HIP kernels")
  if status != OTTX_OK:
    raise OttxError(fn, status)


def require_device() -> torch.device:
  """The GPU the kernels run on; raise loudly when none is present."""
  if not torch.cuda.is_available():
    raise RuntimeError("ott_amd needs a ROCm GPU: torch.cuda.is_available() is False")
  return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> ctypes.c_void_p:
  return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
  return torch.empty((max(int(nbytes), 256),), dtype=torch.uint8, device=device)


def ptr(t: Optional[torch.Tensor]) -> Optional[int]:
  return None if t is None else t.data_ptr()
