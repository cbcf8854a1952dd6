"""sphgpu_mlp_gradient_device (the entry point that allocates nothing; its
scratch is a static device array) against sphgpu_mlp_gradient on the same
network and data: bit-identical gradients and losses, on two calls in a row,
the second with a smaller batch, and the size limit.  Device buffers for the
inputs and outputs come from the HIP runtime the library itself links."""
import ctypes as C

import numpy as np
import pytest

from graphics_amd import sphgpu as sg

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class _Dev:
    """One device buffer holding a copy of a host array."""

    def __init__(self, hip, host: np.ndarray):
        self.hip, self.nbytes, self.p = hip, host.nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.nbytes, 256)) == 0
        assert hip.hipMemcpy(self.p, host.ctypes.data, self.nbytes, _H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    feat = rng.normal(size=(n, 8)).astype(np.float32)
    tgt = (np.tanh(feat[:, 0]) + 0.3 * feat[:, 1] * feat[:, 4]).astype(np.float32)
    return feat, tgt


def test_gradient_device_matches_gradient(gpu):
    hip = _hip()
    lib = sg.load()
    m = sg.Mlp(4096, seed=11)
    rng = np.random.default_rng(3)
    params = (m.get_params() + rng.normal(scale=0.05, size=sg.MLP_PARAMS)).astype(np.float32)
    m.set_params(params)
    dparams = _Dev(hip, params)
    dgrad = _Dev(hip, np.zeros(sg.MLP_PARAMS, np.float32))
    dloss = _Dev(hip, np.zeros(1, np.float64))
    # second call smaller, with a different loss weight
    for n, over, seed in ((4096, 1.0, 1), (1000, 2.5, 2)):
        print(f"[no_alloc] n={n} over_weight={over}", flush=True)
        feat, tgt = _data(n, seed)
        m.set_loss(over)
        g_ref, l_ref = m.gradient(feat, tgt)
        dfeat, dtgt = _Dev(hip, feat), _Dev(hip, tgt)
        rc = lib.sphgpu_mlp_gradient_device(dparams.p, n, dfeat.p, dtgt.p, over, dgrad.p, dloss.p, None)
        assert rc == 0
        assert hip.hipDeviceSynchronize() == 0
        g = dgrad.get(np.zeros(sg.MLP_PARAMS, np.float32))
        loss = dloss.get(np.zeros(1, np.float64))[0]
        assert g.tobytes() == np.asarray(g_ref, np.float32).tobytes()
        assert loss == l_ref
        dfeat.free()
        dtgt.free()
    # beyond the static scratch: refused before any work is queued
    assert lib.sphgpu_mlp_gradient_device(dparams.p, 4097, dparams.p, dparams.p, 1.0, dgrad.p, dloss.p,
                                          None) == -4   # SPHGPU_E_CAPACITY
    assert lib.sphgpu_mlp_gradient_device(None, 10, dparams.p, dparams.p, 1.0, dgrad.p, dloss.p, None) == -1   # SPHGPU_E_ARG
    for b in (dparams, dgrad, dloss):
        b.free()
    m.close()
