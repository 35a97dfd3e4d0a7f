"""GPU: the BF16 MFMA surrogate network (forward, backward, Adam, training)
against the CPU FP64 oracle.  With bf16=True the oracle rounds exactly
where the GPU stores BF16, so the two differ by accumulation order only;
against the unrounded FP64 network the gap is the BF16 rounding itself."""
import numpy as np
import pytest

from cudadsmc_amd.mlp import MLP
from oracle import mlp_oracle as O

pytestmark = pytest.mark.gpu

WIDE = [7, 256, 256, 256, 256, 4]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _layers(widths, g):
    out, o = [], 0
    for i, j in zip(widths[:-1], widths[1:]):
        out.append((g[o:o + i * j], g[o + i * j:o + i * j + j]))
        o += i * j + j
    return out


def test_forward_matches_oracle(gpu_lib):
    rng = np.random.default_rng(1)
    p = O.init_params(WIDE, seed=2) + rng.normal(0, 0.02, sum(i * j + j for i, j in zip(WIDE[:-1], WIDE[1:]))).astype(np.float32)
    net = MLP(WIDE, max_batch=4096, params=p)
    x = rng.normal(size=(5000, 7)).astype(np.float32)            # two batches, the second partial
    y = net.forward(x)
    yb = O.forward(WIDE, p, x, bf16=True)
    yx = O.forward(WIDE, p, x, bf16=False)
    print(f"[mlp] forward rel L2 vs bf16 oracle {_rel(y, yb):.2e}, vs fp64 {_rel(y, yx):.2e}", flush=True)
    # measured 1.5e-4 against the BF16-rounding oracle (accumulation order
    # only) and 4.5e-3 against the unrounded network
    assert _rel(y, yb) < 2e-3
    assert _rel(y, yx) < 2e-2
    net.close()


@pytest.mark.parametrize("batch", [2048, 65536])
def test_gradient_matches_oracle(gpu_lib, batch):
    """Loss and every layer's weight and bias gradient against the
    BF16-rounding FP64 oracle; batch 65,536 is the training batch, whose
    weight gradients are summed over 64 split-K slices of 1,024 rows."""
    rng = np.random.default_rng(3)
    p = O.init_params(WIDE, seed=4)
    p = p + rng.normal(0, 0.02, p.size).astype(np.float32)
    net = MLP(WIDE, max_batch=batch, params=p)
    x = rng.normal(size=(batch, 7)).astype(np.float32)
    t = rng.normal(size=(batch, 4)).astype(np.float32)
    print(f"[mlp] gradient at batch {batch}", flush=True)
    loss, g = net.gradient(x, t)
    lo, go = O.loss_and_grad(WIDE, p, x, t, bf16=True)
    assert abs(loss - lo) / lo < 1e-3
    # measured at most 2.5e-4 per layer
    for l, ((gw, gb), (ow, ob)) in enumerate(zip(_layers(WIDE, g), _layers(WIDE, go))):
        print(f"[mlp] layer {l}: dW rel {_rel(gw, ow):.2e}, db rel {_rel(gb, ob):.2e}", flush=True)
        assert _rel(gw, ow) < 3e-3, l
        assert _rel(gb, ob) < 3e-3, l
    net.close()


def test_tiny_gradient_vs_finite_differences(gpu_lib):
    """Tiny net: GPU gradient vs central differences of the exact FP64 loss."""
    widths = [7, 16, 12, 4]
    rng = np.random.default_rng(5)
    p = O.init_params(widths, seed=6) + rng.normal(0, 0.05, 7 * 16 + 16 + 16 * 12 + 12 + 12 * 4 + 4).astype(np.float32)
    net = MLP(widths, max_batch=64, params=p)
    x = rng.normal(size=(64, 7)).astype(np.float32)
    t = rng.normal(size=(64, 4)).astype(np.float32)
    _, g = net.gradient(x, t)
    pd = p.astype(np.float64)
    fd = np.empty_like(pd)
    for k in range(pd.size):
        e = np.zeros_like(pd)
        e[k] = 1e-6
        fd[k] = (O.loss_and_grad(widths, pd + e, x, t)[0] - O.loss_and_grad(widths, pd - e, x, t)[0]) / 2e-6
    print(f"[mlp] tiny gradient rel L2 vs finite differences {_rel(g, fd):.2e}", flush=True)
    assert _rel(g, fd) < 3e-2
    net.close()


def test_adam_step_matches_closed_form(gpu_lib):
    widths = [7, 64, 4]
    rng = np.random.default_rng(7)
    p = O.init_params(widths, seed=8)
    net = MLP(widths, max_batch=64, params=p)
    q, m, v = p.astype(np.float64), np.zeros(p.size), np.zeros(p.size)
    for t in range(1, 4):
        g = rng.normal(size=p.size).astype(np.float32)
        net.adam(g, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8)
        q, m, v = O.adam(q, g.astype(np.float64), m, v, t, 1e-2, 0.9, 0.999, 1e-8)
    np.testing.assert_allclose(net.params(), q, rtol=1e-4, atol=1e-6)
    net.close()


def test_training_fits_a_smooth_field(gpu_lib):
    """Minibatch Adam at the training batch of 65,536 on a device-resident
    set lowers the loss of a smooth 4-output function of 7 inputs by more
    than 20x."""
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(1 << 20, 7)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 2], np.cos(x[:, 1]) + 0.5 * x[:, 3], x[:, 4] * x[:, 5],
                  np.tanh(x[:, 6] + x[:, 0])], 1).astype(np.float32)
    net = MLP(WIDE, max_batch=65536, seed=10)
    net.set_data(x, t)
    losses = net.train(300, batch=65536, lr=2e-3, seed=11)
    ms, flop = net.train_stats()
    print(f"[mlp] loss {losses[0]:.4f} -> {losses[-1]:.5f}; {ms:.3f} ms/step, {flop / ms / 1e9:.1f} TFLOP/s", flush=True)
    assert losses[-1] < losses[0] / 20
    yh = net.forward(x[:4096])
    assert np.mean((yh - t[:4096]) ** 2) < losses[0] / 20
    net.close()


def test_set_data_twice_then_destroy(gpu_lib):
    """Replacing the training set frees the old one once: a second, smaller
    set trains, and destroying the model afterwards frees each buffer once."""
    rng = np.random.default_rng(12)
    net = MLP(WIDE, max_batch=1024, seed=13)
    for n in (50000, 20000):
        x = rng.uniform(-1, 1, size=(n, 7)).astype(np.float32)
        t = x[:, :4] * 0.5
        net.set_data(x, t)
        losses = net.train(20, batch=1024, lr=1e-3, seed=n)
        assert np.all(np.isfinite(losses))
    net.close()
    again = MLP(WIDE, max_batch=1024, seed=14)      # the allocator is still sound
    again.set_data(x, t)
    assert np.all(np.isfinite(again.train(5, batch=1024)))
    again.close()
