"""Host logic of the differentiable T-step map (no GPU): the fixed-dt guard and the choice
between recording the whole trajectory and checkpointing."""
from types import SimpleNamespace

import pytest


def _box(adaptive):
    return SimpleNamespace(params=SimpleNamespace(adaptive_dt=adaptive))


def test_step_map_refuses_boxes_with_the_cfl_controller():
    """Box() defaults to adaptive_dt = 1; the step adjoints differentiate fixed-dt steps, so the
    T-step map (single box or group) refuses such a box instead of returning a wrong gradient."""
    from dnsbox_amd.adjoint import _require_fixed_dt
    _require_fixed_dt([_box(0), _box(0)])
    with pytest.raises(ValueError, match="adaptive_dt=0"):
        _require_fixed_dt([_box(1)])
    with pytest.raises(ValueError, match="adaptive_dt=0"):
        _require_fixed_dt([_box(0), _box(1)])


def test_segment_length_policy():
    import torch
    from dnsbox_amd.adjoint import _segment_length
    cpu = torch.device("cpu")
    assert _segment_length(200, 20, 1 << 20, cpu, "checkpoint") == 20
    assert _segment_length(200, 20, 1 << 20, cpu, "all") == 200
    assert _segment_length(1, 20, 1 << 20, cpu, "auto") == 20     # one step: nothing to record
    with pytest.raises(ValueError):
        _segment_length(200, 20, 1 << 20, cpu, "sometimes")
    with pytest.raises(ValueError):
        _segment_length(200, 0, 1 << 20, cpu, "checkpoint")
