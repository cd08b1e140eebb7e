"""Argument checks of the ctypes boundary (pyglasstools_amd/engine.py), on
the CPU: device arguments that are not contiguous GPU tensors of the right
dtype and size raise ValueError before anything reaches the library (the C
side cannot check sizes, so such a buffer would be read or written out of
bounds on the device)."""
import numpy as np
import pytest
import torch

from pyglasstools_amd import engine


def fake_engine(dim=2, G=16):
    """An Engine object without a device engine behind it (pgt_ik_create
    needs a GPU); every call below must fail before using the handle."""
    e = engine.Engine.__new__(engine.Engine)
    e.h = None
    e.dim = dim
    e.G = G
    return e


def test_dev_accepts_none_and_rejects_host_memory():
    assert engine._dev(None, torch.float64, 10, "x") is None
    with pytest.raises(ValueError, match="on the GPU"):
        engine._dev(torch.zeros(10, dtype=torch.float64), torch.float64, 10, "x")
    with pytest.raises(ValueError, match="torch tensor"):
        engine._dev(np.zeros(10), torch.float64, 10, "x")
    with pytest.raises(ValueError, match="torch tensor"):
        engine._dev(12345678, torch.float64, 10, "x")


@pytest.mark.parametrize("method", ["compute_frames", "enqueue_frames", "compute_frames_noalloc"])
def test_frame_calls_reject_bad_device_arguments(method):
    e = fake_engine()
    nf, n = 2, 5
    pos = torch.zeros((nf, n, 3), dtype=torch.float64)  # host memory
    diam = torch.ones((nf, n), dtype=torch.float64)
    out = torch.zeros((nf, 16, 2, 2), dtype=torch.float64)
    call = getattr(e, method)
    with pytest.raises(ValueError):
        call(nf, n, pos, diam, fields=("virial",), d_virial=out)
    with pytest.raises(ValueError, match="required"):
        call(nf, n, None, diam, fields=("virial",), d_virial=out)
    with pytest.raises(ValueError, match="non-negative"):
        call(-1, n, pos, diam, fields=("virial",), d_virial=out)


def test_frame_args_sizes_and_dtypes_checked_in_order():
    """_frame_args checks dtype, contiguity and element counts (the device
    check comes first on a CPU tensor, so exercise the rest through a stand-in
    that reports itself on the GPU)."""

    class FakeCuda(torch.Tensor):
        @property
        def is_cuda(self):
            return True

    def fc(t):
        return t.as_subclass(FakeCuda)

    e = fake_engine(dim=2, G=16)
    nf, n = 2, 5
    pos = fc(torch.zeros((nf, n, 3), dtype=torch.float64))
    diam = fc(torch.ones((nf, n), dtype=torch.float64))
    out = fc(torch.zeros((nf, 16, 2, 2), dtype=torch.float64))
    # accepted: returns the mask and pointers
    _, _, mask, ptrs, outs = e._frame_args(nf, n, pos, diam, None, None, None, ("virial",), out, None, None)
    assert mask == engine.FIELD_VIRIAL and ptrs[0] is not None and outs[0] is not None and outs[1] is None
    with pytest.raises(ValueError, match="float64"):
        e._frame_args(nf, n, fc(pos.float()), diam, None, None, None, ("virial",), out, None, None)
    with pytest.raises(ValueError, match="contiguous"):
        e._frame_args(nf, n, fc(torch.zeros((nf, 3, n), dtype=torch.float64)).transpose(1, 2), diam, None, None,
                      None, ("virial",), out, None, None)
    with pytest.raises(ValueError, match="elements"):
        e._frame_args(nf, n + 1, pos, diam, None, None, None, ("virial",), out, None, None)
    with pytest.raises(ValueError, match="elements"):
        e._frame_args(nf, n, pos, diam, None, None, None, ("virial",), fc(out[:1]), None, None)
    with pytest.raises(ValueError, match="required"):
        e._frame_args(nf, n, pos, diam, None, None, None, ("virial", "kinetic"), out, None, None)


def test_network_calls_reject_host_tensors_and_bad_rows():
    k = engine.mlp_nparam()
    p = torch.zeros(k, dtype=torch.float32)
    x = torch.zeros((4, engine.MLP_IN_PAD), dtype=torch.float32)
    y = torch.zeros(4, dtype=torch.float32)
    with pytest.raises(ValueError):
        engine.mlp_infer(p, x, y)
    with pytest.raises(ValueError, match="rows of"):
        engine.mlp_infer(p, torch.zeros((4, 5), dtype=torch.float32), y)
    with pytest.raises(ValueError):
        engine.mlp_adam(p, p, p, p, 1, 1e-3)
