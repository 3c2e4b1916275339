"""Failure handling of batch streams, driven through the fault-injection
build (libtacuda_amd_faults.so: ct_test_stream_batch_fail_next makes the
next operation at a named site report failure; every kernel still runs
normally).  The rules follow the reference's
tests/cpp/test_stream_batch_faults.cpp:
- a failure before the kernel (state upload, staging allocation, input
  upload) leaves the state untouched, so the append can be retried;
- a failure from the launch on (kernel, output download) refuses further
  work, on the GPU and on the CPU, until reset;
- a failed state download when moving to the CPU does not continue there.
Outputs after recovery equal an uninterrupted run and the oracle."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERIES, BARS, PERIOD = 4, 12, 5
GPU, CPU = 2, 1
STATE_UPLOAD, IO_ALLOC, INPUT_UPLOAD, KERNEL, OUTPUT_DOWNLOAD, STATE_DOWNLOAD = 1, 2, 3, 4, 5, 6


@pytest.fixture(scope="module")
def L(dev):
    path = os.path.join(HERE, "tacuda_amd", "libtacuda_amd_faults.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp, i = ctypes.c_void_p, ctypes.c_int
    lib.ct_sma_stream_batch_create.argtypes = [i, i, ctypes.POINTER(vp)]
    lib.ct_sma_stream_batch_append.argtypes = [vp, vp, i, vp]
    lib.ct_stream_batch_set_exec.argtypes = [vp, i]
    lib.ct_stream_batch_reset.argtypes = [vp]
    lib.ct_stream_batch_destroy.argtypes = [vp]
    lib.ct_stream_batch_bars.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    lib.ct_test_stream_batch_fail_next.argtypes = [i]
    lib.ct_test_stream_batch_fail_next.restype = None
    return lib


class Batch:
    def __init__(self, L):
        self.L = L
        self.h = ctypes.c_void_p()
        assert L.ct_sma_stream_batch_create(PERIOD, SERIES, ctypes.byref(self.h)) == 0

    def exec(self, mode):
        return self.L.ct_stream_batch_set_exec(self.h, mode)

    def append(self, x):
        out = np.full_like(x, -1.0)
        rc = self.L.ct_sma_stream_batch_append(self.h, x.ctypes.data, x.shape[1], out.ctypes.data)
        return rc, out

    def bars(self):
        v = ctypes.c_longlong(-1)
        assert self.L.ct_stream_batch_bars(self.h, ctypes.byref(v)) == 0
        return v.value

    def close(self):
        self.L.ct_stream_batch_destroy(self.h)


def matrix():
    s = np.arange(SERIES, dtype=np.float32)[:, None]
    b = np.arange(BARS, dtype=np.float32)[None, :]
    return np.ascontiguousarray(100.0 + 10.0 * s + b, dtype=np.float32)


def same(a, b):
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32))


def uninterrupted(L):
    b = Batch(L)
    assert b.exec(GPU) == 0
    rc, out = b.append(matrix())
    assert rc == 0
    b.close()
    from oracle import ORC

    x = matrix()
    for r in range(SERIES):
        same(out[r], ORC.sma(x[r], PERIOD))
    return out


def test_failures_before_the_kernel_can_be_retried(L):
    expected = uninterrupted(L)
    for site in (STATE_UPLOAD, IO_ALLOC, INPUT_UPLOAD):
        print(f"fault site {site}", flush=True)
        b = Batch(L)
        assert b.exec(GPU) == 0
        L.ct_test_stream_batch_fail_next(site)
        rc, _ = b.append(matrix())
        assert rc != 0, site
        assert b.bars() == 0, site
        rc, out = b.append(matrix())
        assert rc == 0, site
        assert b.bars() == BARS
        same(out, expected)
        b.close()


def test_failures_after_the_launch_refuse_further_work(L):
    expected = uninterrupted(L)
    for site in (KERNEL, OUTPUT_DOWNLOAD):
        print(f"fault site {site}", flush=True)
        b = Batch(L)
        assert b.exec(GPU) == 0
        assert b.append(matrix())[0] == 0
        L.ct_test_stream_batch_fail_next(site)
        assert b.append(matrix())[0] != 0, site
        assert b.bars() == BARS, "a failed append must not count"
        assert b.append(matrix())[0] == 5, site  # CT_STATUS_DEVICE_ERROR
        assert b.exec(CPU) == 0
        assert b.append(matrix())[0] == 5, site
        assert b.bars() == BARS
        assert L.ct_stream_batch_reset(b.h) == 0
        assert b.exec(GPU) == 0
        rc, out = b.append(matrix())
        assert rc == 0
        same(out, expected)
        b.close()


def test_failed_state_download_does_not_continue_on_cpu(L):
    print("fault state download", flush=True)
    b = Batch(L)
    assert b.exec(GPU) == 0
    assert b.append(matrix())[0] == 0
    assert b.exec(CPU) == 0
    L.ct_test_stream_batch_fail_next(STATE_DOWNLOAD)
    assert b.append(matrix())[0] == 2  # CT_STATUS_COPY_FAILED
    assert b.append(matrix())[0] == 5  # CT_STATUS_DEVICE_ERROR
    assert b.bars() == BARS
    assert L.ct_stream_batch_reset(b.h) == 0
    rc, out = b.append(matrix())
    assert rc == 0
    assert b.bars() == BARS
    same(out, uninterrupted(L))
    b.close()
