"""jh_lws_predict_device, the entry point that allocates nothing: its outputs equal jh_lws_predict's
(same kernels, same inputs: bit for bit) on two calls in a row, the second with fewer rows, so the
shared staging array's contents from the first call do not leak into the second; sizes above the
staging maximum and host pointers are rejected before any kernel runs.  The test itself allocates the
device buffers the caller would own."""
import ctypes

import numpy as np
import pytest

from jitterphysics2_amd._abi import LWS_PREDICT_MAX_ROWS, ptr
from jitterphysics2_amd.lws import LearnedWarmStart

pytestmark = pytest.mark.gpu

JH_ERR_INVALID_ARG, JH_ERR_CAPACITY = -1, -3


class DeviceBuffer:
    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, ptr(a), a.nbytes, 1) == 0

    def download(self, a):
        assert self.hip.hipMemcpy(ptr(a), self.p, a.nbytes, 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _model():
    rng = np.random.default_rng(31)
    m = LearnedWarmStart(seed=0x77)
    m.set_normalisation(rng.normal(size=24).astype(np.float32) * 0.1, (0.5 + rng.random(24)).astype(np.float32),
                        np.array([0.02, 0.01, 0.01], np.float32))
    return m, rng


def test_no_alloc_predict_equals_predict_on_two_calls():
    m, rng = _model()
    fbuf = DeviceBuffer(LWS_PREDICT_MAX_ROWS * 24 * 4)
    obuf = DeviceBuffer(LWS_PREDICT_MAX_ROWS * 3 * 4)
    try:
        for n in (40000, 1237):  # the second call is smaller
            f = rng.normal(size=(n, 24)).astype(np.float32)
            want = m.predict(f)
            fbuf.upload(f)
            print(f"no-alloc predict: {n} rows", flush=True)
            st = m.lib.jh_lws_predict_device(m.handle, fbuf.p, n, obuf.p)
            assert st == 0, m.lib.jh_last_error()
            got = obuf.download(np.zeros((n, 3), np.float32))
            assert np.array_equal(got, want), np.abs(got - want).max()
        # the largest size the staging array holds
        n = LWS_PREDICT_MAX_ROWS
        f = rng.normal(size=(n, 24)).astype(np.float32)
        fbuf.upload(f)
        assert m.lib.jh_lws_predict_device(m.handle, fbuf.p, n, obuf.p) == 0
        assert np.array_equal(obuf.download(np.zeros((n, 3), np.float32)), m.predict(f))
    finally:
        fbuf.free()
        obuf.free()


def test_no_alloc_predict_rejects_oversize_and_host_pointers():
    m, rng = _model()
    fbuf = DeviceBuffer(1024 * 24 * 4)
    obuf = DeviceBuffer(1024 * 3 * 4)
    try:
        assert m.lib.jh_lws_predict_device(m.handle, fbuf.p, LWS_PREDICT_MAX_ROWS + 1, obuf.p) == JH_ERR_CAPACITY
        assert b"staging" in m.lib.jh_last_error()
        host = np.zeros((16, 24), np.float32)
        assert m.lib.jh_lws_predict_device(m.handle, ptr(host), 16, obuf.p) == JH_ERR_INVALID_ARG
        assert m.lib.jh_lws_predict_device(m.handle, fbuf.p, 16, ptr(host)) == JH_ERR_INVALID_ARG
        assert m.lib.jh_lws_predict_device(m.handle, fbuf.p, 0, obuf.p) == JH_ERR_INVALID_ARG
        # and the entry point still works afterwards
        f = rng.normal(size=(16, 24)).astype(np.float32)
        fbuf.upload(f)
        assert m.lib.jh_lws_predict_device(m.handle, fbuf.p, 16, obuf.p) == 0
        assert np.array_equal(obuf.download(np.zeros((16, 3), np.float32)), m.predict(f))
    finally:
        fbuf.free()
        obuf.free()
