"""CPU checks of the Python binding's argument validation: mismatched input
shapes, ragged inputs whose length differs from the layout's last offset,
and streaming appends whose arrays disagree are rejected with ValueError
before anything reaches the library (the reference binding rejects
mismatched lengths too, bindings/python/tacuda/__init__.py:852,914)."""
import numpy as np
import pytest
import torch

from tacuda_amd import ops
from tacuda_amd.stream import Stream, StreamBatch


def test_mismatched_hlc_shapes_are_rejected():
    h = torch.zeros(100)
    l = torch.zeros(99)
    c = torch.zeros(100)
    with pytest.raises(ValueError, match="shapes differ"):
        ops.atr(h, l, c, 14)
    with pytest.raises(ValueError, match="shapes differ"):
        ops.stochastic(h, c, torch.zeros(4, 25), 14, 3)
    with pytest.raises(ValueError, match="shapes differ"):
        ops.pack(h, l, c)
    with pytest.raises(ValueError, match="shapes differ"):
        ops.cdl_all(h, h, l, c)


def test_cpu_tensors_are_rejected():
    x = torch.zeros(100)
    with pytest.raises(ValueError, match="on the GPU"):
        ops.ema(x, 20)
    with pytest.raises(ValueError, match="2-D"):
        ops.ema(torch.zeros(2, 3, 4), 20)


class _FakeLayout:
    offsets = np.array([0, 10, 30], dtype=np.int64)


def test_ragged_length_must_match_offsets():
    lay = _FakeLayout()
    with pytest.raises(ValueError, match="offsets"):
        ops._check_ragged(lay, torch.zeros(29))
    with pytest.raises(ValueError, match="offsets"):
        ops._check_ragged(lay, torch.zeros(30), torch.zeros(31), torch.zeros(30))
    with pytest.raises(ValueError, match="offsets"):
        ops._check_ragged(lay, torch.zeros(3, 10))
    # the right length passes the length check and stops at the device check
    with pytest.raises(ValueError, match="on the GPU"):
        ops._check_ragged(lay, torch.zeros(30))


def test_stream_append_rejects_short_inputs():
    s = Stream("atr", 14)
    with pytest.raises(ValueError, match="shapes differ"):
        s.append(high=np.ones(10), low=np.ones(9), close=np.ones(10))
    with pytest.raises(ValueError, match="missing"):
        s.append(close=np.ones(10))
    assert s.bars == 0
    s.close()


def test_stream_batch_append_rejects_bad_shapes():
    b = StreamBatch("atr", 4, 14, exec="cpu")
    with pytest.raises(ValueError, match="shapes differ"):
        b.append(high=np.ones((4, 5)), low=np.ones((4, 4)), close=np.ones((4, 5)))
    with pytest.raises(ValueError, match="split"):
        b.append(high=np.ones(10), low=np.ones(10), close=np.ones(10))
    assert b.bars == 0
    b.close()


def test_stream_batch_append_d_rejects_bad_tensors():
    b = StreamBatch("ema", 4, 20, exec="cpu")
    with pytest.raises(ValueError, match="split"):
        b.append_d(close=torch.zeros(10))
    with pytest.raises(ValueError, match="missing"):
        b.append_d(high=torch.zeros(8))
    with pytest.raises(ValueError, match="one device"):
        b.append_d(close=torch.zeros(8), outs=[torch.zeros(4)])
    with pytest.raises(ValueError, match="on the GPU"):
        b.append_d(close=torch.zeros(8))
    b.close()
