"""CPU: the FP64 oracle of the surrogate network against central finite
differences (its backward pass) and closed forms (BF16 rounding, Adam)."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_closed_forms():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -10, -3.0, 0.0], np.float32)
    r = M.round_bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even (down)
    assert r[2] == 1.0 + 2 ** -6            # tie -> even (up)
    assert r[3] == 1.0 + 2 ** -7            # above the tie -> up
    assert r[4] == -3.0 and r[5] == 0.0
    rng = np.random.default_rng(0)
    v = rng.normal(size=10000).astype(np.float32)
    rel = np.abs(M.round_bf16(v) - v) / np.abs(v)
    assert rel.max() <= 2.0 ** -8 + 1e-12


def test_backward_matches_finite_differences():
    widths = [7, 16, 12, 4]
    p = M.init_params(widths, seed=3).astype(np.float64)
    p += np.random.default_rng(4).normal(0, 0.05, p.size)     # nonzero biases too
    rng = np.random.default_rng(5)
    x = rng.normal(size=(24, 7))
    t = rng.normal(size=(24, 4))
    loss, g = M.loss_and_grad(widths, p, x, t)
    h = 1e-6
    fd = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (M.loss_and_grad(widths, p + e, x, t)[0] - M.loss_and_grad(widths, p - e, x, t)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9)
    assert loss == np.mean((M.forward(widths, p, x) - t) ** 2)


def test_adam_closed_form_first_step():
    """First bias-corrected step moves every parameter by lr * sign(g)
    (up to eps): m_hat = g, v_hat = g^2."""
    p = np.zeros(5)
    g = np.array([1.0, -2.0, 0.5, 3.0, -0.25])
    q, m, v = M.adam(p, g, np.zeros(5), np.zeros(5), 1, 0.01, 0.9, 0.999, 1e-12)
    np.testing.assert_allclose(q, -0.01 * np.sign(g), rtol=1e-9)


def test_package_initialisation_matches_the_oracle_layout():
    """The package's default weights (cudadsmc_amd.mlp.init_params) have the
    oracle's flat layout and Glorot bounds: every weight within
    sqrt(6 / (fan_in + fan_out)) of zero, every bias zero."""
    from cudadsmc_amd import mlp as P
    widths = [7, 16, 12, 4]
    p = P.init_params(widths, seed=3)
    assert p.dtype == np.float32 and p.shape == M.init_params(widths, seed=3).shape
    layers = M.unflatten(widths, p.astype(np.float64))
    for (W, b), i, j in zip(layers, widths[:-1], widths[1:]):
        assert W.shape == (j, i)
        assert np.all(np.abs(W) <= np.sqrt(6.0 / (i + j)))
        assert np.all(b == 0.0)
    assert not np.array_equal(P.init_params(widths, seed=3), P.init_params(widths, seed=4))
