"""Option validation of the host layer that runs before any GPU call: options
the HIP path does not run raise NotImplementedError instead of being ignored."""

import numpy as np
import pytest
import torch

from ott_amd.solvers.linear import acceleration, sinkhorn
from ott_amd.tools import segment_sinkhorn


def test_adaptive_momentum_limits_on_the_host():
  ok = sinkhorn.Sinkhorn(inner_iterations=5, momentum=acceleration.Momentum(start=10))
  ok._check_options()
  assert ok.momentum.start == 10 and ok.momentum.inner_iterations == 5
  with pytest.raises(NotImplementedError):  # two recorded blocks must precede start
    sinkhorn.Sinkhorn(inner_iterations=5, momentum=acceleration.Momentum(start=9))._check_options()
  sinkhorn.Sinkhorn(inner_iterations=1, norm_error=2, momentum=acceleration.Momentum(start=4))._check_options()
  with pytest.raises(NotImplementedError):
    sinkhorn.Sinkhorn(inner_iterations=1, parallel_dual_updates=True,
                      momentum=acceleration.Momentum(start=4))._check_options()


def test_batched_solve_refuses_adaptive_momentum_and_schedules():
  x = torch.zeros((8, 2))
  with pytest.raises(NotImplementedError):
    segment_sinkhorn.segment_sinkhorn(x, x, num_segments=2, max_measure_size=4, epsilon=0.1,
                                      sinkhorn_kwargs=dict(momentum=acceleration.Momentum(start=20)))
  from ott_amd.geometry import epsilon_scheduler
  with pytest.raises(NotImplementedError):
    segment_sinkhorn.segment_sinkhorn(x, x, num_segments=2, max_measure_size=4,
                                      epsilon=epsilon_scheduler.Epsilon(0.1, init=5.0, decay=0.9))


def test_abi_struct_offsets():
  """The ctypes mirror of ottx_sinkhorn_opts keeps the header's layout."""
  from ott_amd import _lib
  o = _lib.SinkhornOpts
  assert o.parallel_dual_updates.offset == 48
  assert o.epsilon_init.offset == 56 and o.epsilon_decay.offset == 64
  assert o.momentum_start.offset == 72 and o.momentum_error_threshold.offset == 80
  assert np.isclose(_lib.SinkhornOpts(momentum_error_threshold=np.inf).momentum_error_threshold, np.inf)


def test_reg_ot_cost_gradient_limits_on_the_host():
  from ott_amd.solvers.linear import danskin
  x = torch.zeros((8, 3))
  with pytest.raises(NotImplementedError):  # a relative epsilon would add terms through eps
    danskin.reg_ot_cost(x, x, epsilon=None)
  with pytest.raises(NotImplementedError):
    danskin.reg_ot_cost(x, x, epsilon=0.1, scale_cost="mean")
  with pytest.raises(NotImplementedError):
    danskin.reg_ot_cost(x, x, epsilon=0.1, tau_a=0.5)
  with pytest.raises(NotImplementedError):
    danskin.reg_ot_cost(torch.zeros((8, 65)), torch.zeros((8, 65)), epsilon=0.1)
