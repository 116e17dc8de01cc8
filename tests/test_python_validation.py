"""CPU checks of the Python view's buffer validation: a device buffer handed to
a kernel must be a contiguous CUDA tensor of the right element type and size,
so that a wrong argument raises instead of faulting the GPU."""
import numpy as np
import pytest
import torch

import atg
from atg import _device_buffer


class FakeCudaTensor:
    """Stands in for a CUDA tensor (no GPU here): the attributes the check reads."""

    def __init__(self, dtype, n, contiguous=True):
        self.is_cuda = True
        self.dtype = dtype
        self._n = n
        self._c = contiguous

    def is_contiguous(self):
        return self._c

    def numel(self):
        return self._n


def test_none_passes():
    assert _device_buffer(None, "i32", 10, "out") is None


@pytest.mark.parametrize("host", [np.zeros(10, np.int32), torch.zeros(10, dtype=torch.int32), [0] * 10])
def test_host_buffers_rejected(host):
    with pytest.raises(atg.AtgError, match="device"):
        _device_buffer(host, "i32", 10, "out")


def test_good_device_buffer_accepted():
    t = FakeCudaTensor(torch.int32, 10)
    assert _device_buffer(t, "i32", 10, "out") is t
    assert _device_buffer(FakeCudaTensor(torch.int32, 12), "i32", 10, "out") is not None
    # histogram counters: int32 or uint32
    assert _device_buffer(FakeCudaTensor(torch.int32, 5), "u32", 5, "hist") is not None


@pytest.mark.parametrize("t,kind,n,exact,msg", [
    (FakeCudaTensor(torch.int64, 10), "i32", 10, False, "dtype"),
    (FakeCudaTensor(torch.float32, 10), "f64", 10, True, "dtype"),
    (FakeCudaTensor(torch.int32, 9), "i32", 10, False, "elements"),
    (FakeCudaTensor(torch.float64, 11), "f64", 10, True, "elements"),
    (FakeCudaTensor(torch.int32, 10, contiguous=False), "i32", 10, False, "contiguous"),
])
def test_bad_device_buffers_rejected(t, kind, n, exact, msg):
    with pytest.raises(atg.AtgError, match=msg):
        _device_buffer(t, kind, n, "buf", exact=exact)
