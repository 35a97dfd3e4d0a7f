"""The surrogate network on the GPU through the C ABI of include/dsmc_mlp.h
(BF16 MFMA forward and backward kernels, FP32 Adam)."""
import ctypes as C

import numpy as np

from . import abi

_SIG = {
    "dsmc_mlp_create": [C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_void_p)],
    "dsmc_mlp_destroy": [C.c_void_p],
    "dsmc_mlp_nparams": [C.c_void_p, C.POINTER(C.c_int64)],
    "dsmc_mlp_set_params": [C.c_void_p, C.POINTER(C.c_float)],
    "dsmc_mlp_get_params": [C.c_void_p, C.POINTER(C.c_float)],
    "dsmc_mlp_forward": [C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)],
    "dsmc_mlp_gradient": [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                          C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "dsmc_mlp_adam": [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float, C.c_float],
    "dsmc_mlp_set_data": [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64],
    "dsmc_mlp_train": [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint64,
                       C.POINTER(C.c_float)],
    "dsmc_mlp_train_stats": [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)],
}


class MLPError(RuntimeError):
    pass


def _lib():
    lib = abi.load()
    if not getattr(lib, "_mlp_sigs", False):
        for name, args in _SIG.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int
        lib.dsmc_mlp_last_error.argtypes = []
        lib.dsmc_mlp_last_error.restype = C.c_char_p
        lib._mlp_sigs = True
    return lib


def _check(lib, rc):
    if rc != 0:
        raise MLPError(f"dsmc_mlp status {rc}: {lib.dsmc_mlp_last_error().decode(errors='replace')}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(widths, seed=0):
    """Glorot-uniform weights and zero biases (float32), in the flat layout
    dsmc_mlp_set_params takes: per layer, W (out x in, row-major), then b."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, j in zip(widths[:-1], widths[1:]):
        a = np.sqrt(6.0 / (i + j))
        parts += [rng.uniform(-a, a, i * j), np.zeros(j)]
    return np.concatenate(parts).astype(np.float32)


class MLP:
    """widths[0] inputs -> hidden widths (SiLU) -> widths[-1] outputs."""

    def __init__(self, widths, max_batch=65536, params=None, seed=0):
        self.lib = _lib()
        self.widths = list(map(int, widths))
        w = (C.c_int * len(self.widths))(*self.widths)
        self.h = C.c_void_p()
        _check(self.lib, self.lib.dsmc_mlp_create(w, len(self.widths) - 1, int(max_batch), C.byref(self.h)))
        n = C.c_int64()
        _check(self.lib, self.lib.dsmc_mlp_nparams(self.h, C.byref(n)))
        self.nparams = n.value
        self.max_batch = int(max_batch)
        self.set_params(init_params(self.widths, seed) if params is None else params)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.dsmc_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        _check(self.lib, self.lib.dsmc_mlp_set_params(self.h, _p(p)))

    def params(self):
        p = np.empty(self.nparams, np.float32)
        _check(self.lib, self.lib.dsmc_mlp_get_params(self.h, _p(p)))
        return p

    def forward(self, x):
        x = _f32(x).reshape(-1, self.widths[0])
        y = np.empty((x.shape[0], self.widths[-1]), np.float32)
        _check(self.lib, self.lib.dsmc_mlp_forward(self.h, _p(x), x.shape[0], _p(y)))
        return y

    def gradient(self, x, t):
        x = _f32(x).reshape(-1, self.widths[0])
        t = _f32(t).reshape(-1, self.widths[-1])
        loss = np.zeros(1, np.float32)
        g = np.empty(self.nparams, np.float32)
        _check(self.lib, self.lib.dsmc_mlp_gradient(self.h, _p(x), _p(t), x.shape[0], _p(loss), _p(g)))
        return float(loss[0]), g

    def adam(self, grad=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        g = None if grad is None else _f32(grad)
        _check(self.lib, self.lib.dsmc_mlp_adam(self.h, _p(g), lr, beta1, beta2, eps))

    def set_data(self, x, t):
        self._x = _f32(x).reshape(-1, self.widths[0])
        self._t = _f32(t).reshape(-1, self.widths[-1])
        _check(self.lib, self.lib.dsmc_mlp_set_data(self.h, _p(self._x), _p(self._t), self._x.shape[0]))

    def train(self, steps, batch=65536, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=1):
        losses = np.zeros(max(int(steps), 1), np.float32)
        _check(self.lib, self.lib.dsmc_mlp_train(self.h, int(steps), int(batch), lr, beta1, beta2, eps,
                                                 int(seed), _p(losses)))
        return losses[:int(steps)]

    def train_stats(self):
        ms, fl = C.c_double(), C.c_double()
        _check(self.lib, self.lib.dsmc_mlp_train_stats(self.h, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value


def _train_rate(steps=100, batch=65536, width=256, depth=4, log=print):
    """Time `steps` training steps of the surrogate's shape (7 -> width x depth
    -> 4) on a synthetic smooth target, after 20 untimed steps."""
    import time
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1 << 20, 7)).astype(np.float32)
    t = np.tanh(x[:, :4] + 0.5 * x[:, 3:7]).astype(np.float32)
    net = MLP([7] + [width] * depth + [4], max_batch=batch, seed=1)
    net.set_data(x, t)
    net.train(20, batch=batch, lr=1e-3, seed=1)
    t0 = time.perf_counter()
    losses = net.train(steps, batch=batch, lr=1e-3, seed=2)
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ms, flop = net.train_stats()
    log(f"train {steps} steps at batch {batch}: {wall:.3f} ms/step wall, {ms:.3f} ms/step (events), "
        f"{flop / ms / 1e9:.1f} TFLOP/s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    net.close()
    return ms, flop


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time the surrogate's training step on the GPU")
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--batch", type=int, default=65536)
    a = ap.parse_args()
    _train_rate(a.steps, a.batch, log=lambda m: print(m, flush=True))
