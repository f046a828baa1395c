"""Binding of the bucket-width MLP (8 -> 32 -> 32 -> 2, fp32, ReLU) whose
forward, backward and Adam kernels run on the GPU (vine_amd/csrc/mlp.hip).

Parameter layout (1,426 floats): W1[32,8] b1[32] W2[32,32] b2[32] W3[2,32]
b3[2] xmu[8] xsd[8]; the last 16 normalise the inputs and are not trained.
"""
import ctypes

import numpy as np

from ._lib import lib

NIN, NH, NOUT = 8, 32, 2
NTRAIN = NH * NIN + NH + NH * NH + NH + NOUT * NH + NOUT
NPARAM = NTRAIN + 2 * NIN

_declared = False


def _L():
    global _declared
    L = lib()
    if not _declared:
        vp, i32, i64, f32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_uint64
        for name, res, args in (
            ("vine_mlp_param_count", i32, []),
            ("vine_mlp_last_error", ctypes.c_char_p, []),
            ("vine_mlp_create", i32, [ctypes.POINTER(vp)]),
            ("vine_mlp_destroy", i32, [vp]),
            ("vine_mlp_set_params", i32, [vp, vp]),
            ("vine_mlp_get_params", i32, [vp, vp]),
            ("vine_mlp_forward", i32, [vp, vp, i64, vp]),
            ("vine_mlp_gradient", i32, [vp, vp, vp, vp, i64, vp, vp]),
            ("vine_mlp_train", i32, [vp, vp, vp, vp, i64, i32, i32, f32, u64, vp]),
        ):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _declared = True
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError("vine_mlp error %d: %s" % (rc, _L().vine_mlp_last_error().decode(errors="replace")))


def init_params(seed=0, xmu=None, xsd=None):
    """He-initialised weights, zero biases, given input normalisation."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, NOUT)):
        p[o:o + fan_in * fan_out] = rng.normal(0, np.sqrt(2.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    p[NTRAIN:NTRAIN + NIN] = 0.0 if xmu is None else xmu
    p[NTRAIN + NIN:] = 1.0 if xsd is None else np.maximum(xsd, 1e-3)
    return p


class MLP:
    def __init__(self, params=None):
        assert _L().vine_mlp_param_count() == NPARAM
        h = ctypes.c_void_p()
        _check(_L().vine_mlp_create(ctypes.byref(h)))
        self.h = h
        self.set_params(init_params() if params is None else params)

    def close(self):
        if getattr(self, "h", None):
            _L().vine_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (NPARAM,)
        _check(_L().vine_mlp_set_params(self.h, p.ctypes.data))

    def params(self):
        p = np.zeros(NPARAM, np.float32)
        _check(_L().vine_mlp_get_params(self.h, p.ctypes.data))
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, NIN)
        y = np.zeros((x.shape[0], NOUT), np.float32)
        _check(_L().vine_mlp_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradient(self, x, t, m):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, NIN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, NOUT)
        m = np.ascontiguousarray(m, np.float32).reshape(-1, NOUT)
        g = np.zeros(NTRAIN, np.float32)
        loss = ctypes.c_double()
        _check(_L().vine_mlp_gradient(self.h, x.ctypes.data, t.ctypes.data, m.ctypes.data, x.shape[0],
                                      g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def train(self, x, t, m, epochs=20, batch=256, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, NIN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, NOUT)
        m = np.ascontiguousarray(m, np.float32).reshape(-1, NOUT)
        curve = np.zeros(epochs, np.float64)
        _check(_L().vine_mlp_train(self.h, x.ctypes.data, t.ctypes.data, m.ctypes.data, x.shape[0], epochs,
                                   batch, lr, seed, curve.ctypes.data))
        return curve
