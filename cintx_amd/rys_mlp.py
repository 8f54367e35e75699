"""Learned Rys-root initializer: host side.

The network (2 inputs, two tanh layers of 64, 26 outputs) is trained and
evaluated by HIP kernels in libcintx_hip.so (csrc/rys_mlp.hip).  This module
holds the parameter layout, a seeded initialisation, a NumPy restatement of
the forward/backward pass (the reference the device gradients are checked
against) and thin wrappers over the C ABI.
"""
import ctypes

import numpy as np

from . import capi

H, IN, OUT = 64, 2, 26
OW1 = 0
OB1 = OW1 + H * IN
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + OUT * H
NPARAM = OB3 + OUT


def _lib():
    L = capi.lib()
    if not getattr(L, "_mlp_ready", False):
        f32p = ctypes.POINTER(ctypes.c_float)
        L.cintx_rys_mlp_nparam.restype = ctypes.c_int
        L.cintx_rys_mlp_train.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_uint64, capi.DP]
        L.cintx_rys_mlp_grad.argtypes = [f32p, f32p, capi.IP, f32p, ctypes.c_int, f32p, f32p]
        L.cintx_rys_mlp_eval.argtypes = [f32p, ctypes.c_int, capi.DP, ctypes.c_int64,
                                         ctypes.c_int, ctypes.c_int, capi.DP, capi.DP, capi.IP,
                                         capi.DP, capi.DP]
        L._mlp_ready = True
    return L


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def features(x, n):
    return np.stack([(np.log(x) + 3.45) / 10.35, (np.asarray(n) - 7) / 6.0], axis=-1)


def init_params(seed=0):
    """Seeded Glorot-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)

    def glorot(fan_in, fan_out):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-lim, lim, size=(fan_out, fan_in)).astype(np.float32)

    p[OW1:OB1] = glorot(IN, H).ravel()
    p[OW2:OB2] = glorot(H, H).ravel()
    p[OW3:OB3] = glorot(H, OUT).ravel()
    return p


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(H, IN), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3].reshape(OUT, H), p[OB3:])


def mask(n):
    n = np.asarray(n)
    r = np.arange(OUT) % 13
    return (r[None, :] < n[:, None]).astype(np.float64)


def loss_np(p, z, n, tgt):
    """per-sample loss sum over valid outputs of (y - t)^2 / (2n)"""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(z @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    d = (y - tgt) * mask(n)
    return (d * d).sum(axis=1) / (2.0 * np.asarray(n))


def grad_np(p, z, n, tgt):
    """batch-mean gradient of loss_np by back-propagation (float64)"""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = np.asarray(n, dtype=np.float64)
    h1 = np.tanh(z @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    dy = 2.0 * (y - tgt) * mask(n) / (2.0 * n[:, None])
    d2 = (dy @ W3) * (1 - h2 * h2)
    d1 = (d2 @ W2) * (1 - h1 * h1)
    B = len(n)
    g = np.zeros(NPARAM)
    g[OW1:OB1] = (d1.T @ z).ravel() / B
    g[OB1:OW2] = d1.sum(0) / B
    g[OW2:OB2] = (d2.T @ h1).ravel() / B
    g[OB2:OW3] = d2.sum(0) / B
    g[OW3:OB3] = (dy.T @ h2).ravel() / B
    g[OB3:] = dy.sum(0) / B
    return g


def device_grad(p, z, n, tgt):
    L = _lib()
    p = np.ascontiguousarray(p, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    n = np.ascontiguousarray(n, dtype=np.int32)
    tgt = np.ascontiguousarray(tgt, dtype=np.float32)
    g = np.zeros(NPARAM, dtype=np.float32)
    loss = np.zeros(len(n), dtype=np.float32)
    capi._check(L.cintx_rys_mlp_grad(_fp(p), _fp(z), capi._ip(n), _fp(tgt), len(n), _fp(g),
                                     _fp(loss)))
    return g, loss


def train(params=None, nsamples=1 << 22, epochs=8, batch=1024, lr=3e-3, seed=12345):
    L = _lib()
    p = init_params(0) if params is None else np.array(params, dtype=np.float32)
    loss = ctypes.c_double()
    capi._check(L.cintx_rys_mlp_train(_fp(p), nsamples, epochs, batch, lr, seed,
                                      ctypes.byref(loss)))
    return p, loss.value


def evaluate(params, n, x, newton_steps=2, init_mode=0, with_error=False):
    """(u, w, status, device_ms[, polish_err]): status 1 = the polished rule
    is within 1e-14 relative of the table rule (and is returned), 0 = table
    fallback; polish_err = that largest relative difference."""
    L = _lib()
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.zeros(len(x) * n)
    w = np.zeros(len(x) * n)
    st = np.zeros(len(x), dtype=np.int32)
    err = np.zeros(len(x))
    ms = ctypes.c_double()
    capi._check(L.cintx_rys_mlp_eval(_fp(p), n, capi._dp(x), len(x), newton_steps, init_mode,
                                     capi._dp(u), capi._dp(w), capi._ip(st), capi._dp(err),
                                     ctypes.byref(ms)))
    if with_error:
        return u.reshape(-1, n), w.reshape(-1, n), st, ms.value, err
    return u.reshape(-1, n), w.reshape(-1, n), st, ms.value
