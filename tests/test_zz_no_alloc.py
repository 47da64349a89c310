"""qcbf_measure_device, the allocation-free measurement, against qcbf_measure
of a handle holding the same field: bit-identical {P, E, Q}, in both storage
forms, on two calls in a row (the second on a smaller lattice, so the shared
scratch holds the first call's partials beyond the second's), and its
rejection of volumes beyond its scratch.  The device fields are made with
the HIP runtime through ctypes (the test's own allocations; the entry point
makes none)."""
import ctypes

import numpy as np
import pytest

import oracle
import qcbflow

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceBuffer:
    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0  # host -> device

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self):
        hip().hipFree(self.p)


def soa(links, precision):
    """The library's device layout of a (V, 4, 3, 3, 2) field."""
    V = links.shape[0]
    if precision == qcbflow.FP64:
        return np.ascontiguousarray(links.reshape(V, 4, 9, 2).transpose(1, 2, 0, 3))
    rows = links[:, :, 0:2, :, :].reshape(V, 4, 3, 4).astype(np.float32)
    return np.ascontiguousarray(rows.transpose(1, 2, 0, 3))


def measure_device(links, L, precision):
    buf = DeviceBuffer(soa(links, precision).nbytes)
    out = DeviceBuffer(3 * 8)
    try:
        buf.put(soa(links, precision))
        dims = (ctypes.c_int * 4)(*L)
        rc = qcbflow.load().qcbf_measure_device(buf.p, dims, int(precision), out.p, None)
        assert rc == 0, qcbflow.load().qcbf_last_error(None)
        assert hip().hipDeviceSynchronize() == 0
        return out.get((3,), np.float64)
    finally:
        buf.free()
        out.free()


@pytest.mark.parametrize("precision", [qcbflow.FP64, qcbflow.FP32_MIXED])
def test_measure_device_equals_measure(precision):
    for L, seed in (((8, 8, 8, 8), 3), ((4, 4, 6, 4), 4)):
        print(f"qcbf_measure_device {L}, precision {precision}", flush=True)
        f = oracle.rk3(oracle.hot_start(L, seed), L, 0.02, 2)
        with qcbflow.GradientFlow(L, 0.02, precision) as gf:
            gf.copy(f)
            P, E, Q = gf.measure()
            held = gf.field()  # exactly the words the handle holds (FP32 rows exact)
        got = measure_device(held, L, precision)
        assert got[0] == P and got[1] == E and got[2] == Q, (got, (P, E, Q))


def test_measure_device_rejects_volumes_beyond_its_scratch():
    buf = DeviceBuffer(1024)
    out = DeviceBuffer(64)
    try:
        dims = (ctypes.c_int * 4)(64, 64, 64, 65)
        rc = qcbflow.load().qcbf_measure_device(buf.p, dims, qcbflow.FP64, out.p, None)
        assert rc == qcbflow.QCBF_ERR_ARG
        assert b"2^24 sites" in qcbflow.load().qcbf_last_error(None)
        rc = qcbflow.load().qcbf_measure_device(None, dims, qcbflow.FP64, out.p, None)
        assert rc == qcbflow.QCBF_ERR_ARG
    finally:
        buf.free()
        out.free()
