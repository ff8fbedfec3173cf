"""msd_net_moment_noalloc, the allocation-free form of ss:netMoment() over spins the
caller holds in device memory, against msd_system_net_moment on the same spins.

Two calls in a row, the second on a smaller lattice: the FP64 sums must be
bit-identical (the same reduction kernels with the same block count and order).
The device buffers are made with the HIP runtime libmsd.so itself links against,
so both sides use one runtime."""
import ctypes as C

import numpy as np
import pytest

import maglua_amd as M
from maglua_amd._lib import MagLuaError, call

pytestmark = pytest.mark.gpu

H2D = 1   # hipMemcpyHostToDevice


def _hip():
    h = C.CDLL("libamdhip64.so.7")   # already loaded by libmsd.so: the same runtime instance
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _to_device(hip, a):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _noalloc(spins, lo, n):
    out = (C.c_double * 8)()
    call("msd_net_moment_noalloc", spins, lo, int(n), out)
    return tuple(out)


def test_net_moment_noalloc_equals_system_form():
    hip = _hip()
    for dims, seed in (((64, 48, 5), 1), ((17, 9, 3), 2)):     # the second call is smaller
        n = dims[0] * dims[1] * dims[2]
        rng = np.random.default_rng(seed)
        s = (rng.standard_normal((3, n)) * rng.uniform(0.5, 2.0, n)).astype(np.float32)
        ss = M.SpinSystem(*dims)
        ss.setSpins(*s.astype(np.float64))          # FP32 values: stored exactly on the device
        want = ss.netMoment()
        dev = _to_device(hip, s)
        try:
            got = _noalloc(dev, None, n)
            print(f"noalloc netMoment {dims}: {got[:4]}", flush=True)
            assert got == want
            # low words: the sums of hi + lo in FP64 (order differs from numpy's: 1e-12 relative)
            lo = (rng.standard_normal((3, n)) * 1e-8).astype(np.float32)
            dlo = _to_device(hip, lo)
            try:
                got_lo = _noalloc(dev, dlo, n)
            finally:
                hip.hipFree(dlo)
            v = s.astype(np.float64) + lo.astype(np.float64)
            ref = list(v.sum(axis=1)) + list((v * v).sum(axis=1))
            for k, r in zip((0, 1, 2, 4, 5, 6), ref):
                assert abs(got_lo[k] - r) <= 1e-12 * max(1.0, np.abs(v).sum()), k
        finally:
            hip.hipFree(dev)


def test_net_moment_noalloc_rejects_bad_sizes():
    hip = _hip()
    M.SpinSystem(4, 4, 1)                          # msd_init has run
    dev = _to_device(hip, np.ones(12, np.float32))
    try:
        for n in (0, -5, (1 << 30) + 1):      # rejected on the host before any launch
            with pytest.raises(MagLuaError):
                _noalloc(dev, None, n)
    finally:
        hip.hipFree(dev)
