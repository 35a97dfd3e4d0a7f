"""CPU: the steady-state detector on synthetic signals with a known settling
time -- exponential approach to a steady field plus noise, sampled in
half-windows as a DSMC run is."""
import math

import numpy as np
import pytest

from cudadsmc_amd.steady import SteadyDetector, density


def _base(C):
    return 1.0 + 0.5 * np.sin(np.linspace(0, 6, C))


def _signal(tau_h, A, sigma, C=4000, windows=120, seed=0, corr=0.0):
    """Half-window fields h_k = base (1 + A exp(-k/tau_h)) + noise of s.d.
    sigma per cell; corr > 0 makes the two halves' noise share a part, as
    correlated samples do."""
    rng = np.random.default_rng(seed)
    base = _base(C)
    out = []
    for w in range(windows):
        common = rng.normal(0, 1, C) * math.sqrt(corr)
        h = [base * (1.0 + A * math.exp(-(2 * w + q) / tau_h)) +
             sigma * (common + rng.normal(0, 1, C) * math.sqrt(1 - corr)) for q in (0, 1)]
        out.append(h)
    return out


def _expected(tau_h, A, sigma, thr, pat, C=4000):
    """First window whose expected chi drops below thr, plus patience - 1."""
    B2 = float(np.mean(_base(C) ** 2))
    h = lambda k: A * math.exp(-k / tau_h)
    s2 = lambda w: B2 * (h(2 * w + 1) - h(2 * w)) ** 2 + 2 * sigma ** 2
    for w in range(1, 100000):
        m, mp = 0.5 * (h(2 * w) + h(2 * w + 1)), 0.5 * (h(2 * w - 2) + h(2 * w - 1))
        d2 = B2 * (m - mp) ** 2 + sigma ** 2
        if math.sqrt(2 * d2 / (0.5 * (s2(w) + s2(w - 1)))) < thr:
            return w + 1 + pat - 1        # windows are counted from 1
    return None


@pytest.mark.parametrize("tau_h,A,sigma", [(10.0, 0.5, 0.01), (20.0, 1.0, 0.02), (6.0, 2.0, 0.005)])
def test_detector_settling_time(tau_h, A, sigma):
    thr, pat = 1.5, 3
    det = SteadyDetector(thr, pat)
    for a, b in _signal(tau_h, A, sigma):
        if det.update(a, b):
            break
    exp = _expected(tau_h, A, sigma, thr, pat)
    assert det.steady_at is not None
    assert abs(det.steady_at - exp) <= 2, (det.steady_at, exp)


def test_pure_noise_is_steady_at_once():
    det = SteadyDetector(1.5, 3)
    for a, b in _signal(1.0, 0.0, 0.02, windows=6):
        det.update(a, b)
    assert det.steady_at == 4
    assert all(abs(c - 1.0) < 0.1 for c in det.history)


def test_noise_scale_comes_from_the_data():
    """Steady signals read as steady whatever their noise level (the scale
    is measured from the half-windows, not from a count model), also when
    the two halves of a window share a fifth of their noise (samples near
    the half boundary are correlated): chi^2 = (1 + c)/(1 - c) = 1.5."""
    for sigma in (0.001, 0.05):
        det = SteadyDetector(1.5, 3)
        for a, b in _signal(1.0, 0.0, sigma, windows=8, corr=0.2):
            det.update(a, b)
        assert det.steady_at is not None
        assert abs(np.mean(det.history) - math.sqrt(1.5)) < 0.1


def test_transient_is_never_steady():
    """A drift of 3 % per half-window against 0.5 % noise never triggers."""
    rng = np.random.default_rng(2)
    det = SteadyDetector(1.5, 3)
    base = _base(3000)
    for w in range(40):
        a, b = (base * (1.0 + 0.03 * (2 * w + q)) + 0.005 * rng.normal(0, 1, 3000) for q in (0, 1))
        assert not det.update(a, b)
    assert min(det.history) > 2.0


def test_density_from_counts():
    n = density(np.array([100.0, 0.0, 50.0]), 10, 2.0, np.array([1.0, 1.0, 0.0]))
    assert n.tolist() == [20.0, 0.0, 0.0]


def test_lag_catches_slow_drift():
    """A drift of 0.3 % of the field per half-window under 1 % noise reads as
    steady window to window (chi about 1.3) but not over a lag of 4 windows."""
    rng = np.random.default_rng(4)
    base = _base(4000)
    lag1, lag4 = SteadyDetector(1.5, 3, lag=1), SteadyDetector(1.5, 3, lag=4)
    for w in range(30):
        a, b = (base * (1.0 + 0.003 * (2 * w + q)) + 0.01 * rng.normal(0, 1, 4000) for q in (0, 1))
        lag1.update(a, b)
        lag4.update(a, b)
    assert lag1.steady_at is not None
    assert lag4.steady_at is None


def test_lag_ramps_up_from_the_first_window():
    """Before lag windows have passed, each window is compared with the
    first one: pure noise is confirmed at window patience + 1 whatever the
    lag, and its chi stays about 1 at every lag used."""
    for lag in (1, 2, 4, 8):
        det = SteadyDetector(1.5, 3, lag=lag)
        for a, b in _signal(1.0, 0.0, 0.02, windows=4, seed=lag):
            det.update(a, b)
        assert det.steady_at == 4, lag
        assert all(abs(c - 1.0) < 0.1 for c in det.history)


def test_ramp_keeps_a_drift_from_passing_early():
    """A drift that one window's lag hides (chi about 1.15 at lag 1) shows at
    lags 2 and 3 of the ramp, so a run drifting from its start is not
    confirmed in its first windows."""
    rng = np.random.default_rng(5)
    base = _base(4000)
    det = SteadyDetector(1.5, 3, lag=4)
    for w in range(12):
        a, b = (base * (1.0 + 0.003 * (2 * w + q)) + 0.01 * rng.normal(0, 1, 4000) for q in (0, 1))
        det.update(a, b)
    assert det.history[0] < 1.5 < det.history[1] < det.history[2]
    assert det.steady_at is None


def _three_field_windows(C, windows, drift, seed=0, sigma=0.01):
    """Half-window density, velocity (2 x C) and temperature fields with
    noise of s.d. sigma; `drift` names the field that relaxes with time
    constant 15 half-windows (the others stay steady)."""
    rng = np.random.default_rng(seed)
    base = _base(C)
    out = []
    for w in range(windows):
        halves = []
        for q in (0, 1):
            k = 2 * w + q
            r = 0.4 * math.exp(-k / 15.0)
            n = base * (1 + (r if drift == "n" else 0)) + sigma * rng.normal(0, 1, C)
            u = np.stack([base * (1 + (r if drift == "u" else 0)), 0.3 * base]) + sigma * rng.normal(0, 1, (2, C))
            T = 2 * base * (1 + (r if drift == "T" else 0)) + sigma * rng.normal(0, 1, C)
            halves.append([n, u, T])
        out.append(halves)
    return out


@pytest.mark.parametrize("drift", ["n", "u", "T"])
def test_each_field_is_watched(drift):
    """A relaxation in any one of density, velocity and temperature, with
    the other two steady, holds the detector back: it confirms well after a
    run with no relaxation at all (a density-only detector would confirm a
    velocity or temperature relaxation as early as the steady run)."""
    at = {}
    for which in ("n", "u", "T", None):
        det = SteadyDetector(1.5, 3, lag=4)
        for a, b in _three_field_windows(3000, 60, which, seed=3):
            if det.update(a, b):
                break
        at[which] = det.steady_at
    assert at[None] is not None and at[drift] is not None
    assert at[drift] > at[None] + 5


def test_count_weights_keep_sparse_cells_from_hiding_a_drift():
    """Velocity noise falls as 1/sqrt(samples): a drift in the well-sampled
    cells is seen under count weights even when a few sparse cells carry
    100 times the noise; unweighted, the sparse cells' noise hides it."""
    rng = np.random.default_rng(4)
    C = 4000
    counts = np.where(np.arange(C) < 200, 1.0, 400.0)
    sig = 0.5 / np.sqrt(counts)
    res = {}
    for weighted in (True, False):
        det = SteadyDetector(1.5, 3, lag=4)
        for w in range(40):
            halves = []
            for q in (0, 1):
                r = 0.02 * (2 * w + q)          # a steady linear drift, never settling
                u = np.stack([1.0 + r * (counts > 1) + sig * rng.normal(0, 1, C), sig * rng.normal(0, 1, C)])
                halves.append([u])
            det.update(*halves, weights=[counts] if weighted else None)
        res[weighted] = det.steady_at
    assert res[True] is None
    assert res[False] is not None
