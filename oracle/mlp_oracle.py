"""CPU FP64 oracle of the surrogate MLP (test infrastructure only; the
product runs the network with the HIP kernels of cudadsmc_amd/csrc/mlp.hip).

Network: widths[0] -> ... -> widths[L], SiLU after every layer but the
last, mean squared error loss.  Parameters flat as in include/dsmc_mlp.h:
per layer W[out][in] then b[out].

bf16=True rounds where the GPU stores BF16 -- inputs, weights,
pre-activations, activations, output gradients and back-propagated
gradients -- and computes everything else in FP64, so GPU and oracle
differ only by accumulation order.  bf16=False is the exact FP64 network,
which the finite-difference checks pin.
"""
import numpy as np

__all__ = ["unflatten", "flatten", "init_params", "round_bf16", "forward", "loss_and_grad", "adam"]


def round_bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(widths, p):
    out, o = [], 0
    for i, j in zip(widths[:-1], widths[1:]):
        W = np.asarray(p[o:o + i * j], np.float64).reshape(j, i)
        o += i * j
        b = np.asarray(p[o:o + j], np.float64)
        o += j
        out.append((W, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init_params(widths, seed=0):
    """Glorot-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, j in zip(widths[:-1], widths[1:]):
        a = np.sqrt(6.0 / (i + j))
        parts += [rng.uniform(-a, a, i * j), np.zeros(j)]
    return np.concatenate(parts).astype(np.float32)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(widths, p, x, bf16=False, keep=False):
    r = round_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(widths, p)
    a = r(x)
    acts, zs = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        if l + 1 < len(layers):
            zs.append(r(z))
            a = r(_silu(z))
            acts.append(a)
        else:
            a = z
    return (a, acts, zs) if keep else a


def loss_and_grad(widths, p, x, t, bf16=False):
    """Mean squared error and its gradient (flat like p)."""
    r = round_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(widths, p)
    y, acts, zs = forward(widths, p, x, bf16=bf16, keep=True)
    n, m = t.shape
    e = y - t
    loss = float(np.mean(e * e))
    d = r(2.0 * e / (n * m))
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (d.T @ acts[l], d.sum(axis=0))
        if l > 0:
            d = r((d @ r(W)) * _dsilu(zs[l - 1]))
    return loss, flatten(grads)


def adam(p, g, m, v, t, lr, b1, b2, eps):
    """One bias-corrected Adam step (float64); returns p, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    lr_t = lr * np.sqrt(1 - b2 ** t) / (1 - b1 ** t)
    return p - lr_t * m / (np.sqrt(v) + eps), m, v
