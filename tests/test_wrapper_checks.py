"""The Python wrappers refuse buffers whose addresses must not reach a kernel
(host arrays, CPU tensors, wrong dtype, shape or layout) before any call into
the library, so these run without a GPU."""
import numpy as np
import pytest
import torch

import randomfield_amd as R
from randomfield_amd import serving as S


class _FakeCtx:
    handle = None
    device = 0


def _plan(nDim=3, Nmc=4):
    plan = object.__new__(R.FieldPlan)        # no device plan: the checks come first
    plan.ctx = _FakeCtx()
    plan.params = R.make_params(nDim=nDim, Nmc=Nmc)
    plan._h = None
    return plan


def test_check_device_refuses_host_buffers():
    with pytest.raises(ValueError, match="device tensor"):
        R._check_device(np.zeros((4, 3)), "x")
    with pytest.raises(ValueError, match="device memory"):
        R._check_device(torch.zeros(4, 3, dtype=torch.float64), "x")
    with pytest.raises(ValueError, match="device tensor"):
        R._check_device(12345, "x")


@pytest.mark.parametrize("x,out,msg", [
    (np.zeros((10, 3)), None, "device tensor"),
    (torch.zeros(10, 3, dtype=torch.float64), None, "device memory"),
])
def test_execute_refuses_host_points(x, out, msg):
    plan = _plan()
    with pytest.raises(ValueError, match=msg):
        plan.execute(x, out)
    with pytest.raises(ValueError, match=msg):
        plan.autotune(x, out)


def test_execute_grid_refuses_host_output():
    plan = _plan(Nmc=2)
    with pytest.raises(ValueError, match="device tensor"):
        plan.execute_grid((0, 0, 0), (1, 1, 1), (4, 3, 2), np.zeros((2, 24), np.float32))
    with pytest.raises(ValueError, match="device memory"):
        plan.execute_grid((0, 0, 0), (1, 1, 1), (4, 3, 2), torch.zeros(2, 24))


def test_server_submit_refuses_cpu_tensor_and_bad_shape():
    srv = object.__new__(S.Server)
    srv.ctx, srv._h, srv._keep = _FakeCtx(), None, {}
    with pytest.raises(ValueError, match="device memory"):
        srv.submit(torch.zeros(10, 3, dtype=torch.float64), 2, 4, (1, 1, 1), 1)
    with pytest.raises(ValueError, match="expected \\(N, 3\\)"):
        srv.submit(np.zeros((10, 2)), 2, 4, (1, 1, 1), 1)
