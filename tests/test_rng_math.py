"""CPU: Philox4x32-10 against published known answers and rocRAND, and the
IEEE-only log / exp / sincos(2 pi u) against double-precision libm."""
import json
import os

import numpy as np

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "philox_vectors.json")


def _vectors(key):
    d = json.load(open(GOLD))[key]
    ctr = np.array([v[0] for v in d], np.uint32)
    k = np.array([v[1] for v in d], np.uint32)
    out = np.array([v[2] for v in d], np.uint32)
    return ctr, k, out


def test_philox_random123_kat():
    ctr, k, out = _vectors("random123_kat")
    np.testing.assert_array_equal(O.philox(ctr, k[:, 0], k[:, 1]), out)


def test_philox_matches_rocrand():
    ctr, k, out = _vectors("rocrand")
    np.testing.assert_array_equal(O.philox(ctr, k[:, 0], k[:, 1]), out)


def test_philox_uniformity():
    n = 1 << 16
    ctr = np.zeros((n, 4), np.uint32)
    ctr[:, 0] = np.arange(n)
    ctr[:, 1] = 3
    w = O.philox(ctr, 12345, 678).ravel()
    u = (w >> 8).astype(np.float64) * 2.0 ** -24
    h, _ = np.histogram(u, bins=64, range=(0, 1))
    e = u.size / 64
    chi2 = ((h - e) ** 2 / e).sum()
    assert chi2 < 120, chi2          # 63 dof; p ~ 1e-5 at 120
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)


def _ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_logf_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([
        (rng.random(200000) * 2.0 ** rng.integers(-30, 30, 200000)).astype(np.float32),
        np.arange(1, 1 << 16, dtype=np.float32) * np.float32(2.0 ** -24),   # the u01o grid
        np.array([1.0, 2.0, 0.5, 1e-30, 3e38], np.float32)])
    x = x[x > 0]
    got = O.logf(x)
    ref = np.log(x.astype(np.float64))
    err = _ulp_err(got, ref)
    assert err.max() <= 4.0, err.max()
    assert O.logf(np.float32([1.0]))[0] == 0.0


def test_expf_accuracy():
    rng = np.random.default_rng(2)
    x = (rng.random(400000) * 170 - 85).astype(np.float32)
    got = O.expf(x)
    ref = np.exp(x.astype(np.float64))
    err = _ulp_err(got, ref)
    assert err.max() <= 4.0, err.max()
    assert O.expf(np.float32([0.0]))[0] == 1.0
    assert O.expf(np.float32([-100.0]))[0] == 0.0


def test_sincos2pi_accuracy():
    u = (np.arange(1 << 20, dtype=np.float64) * 2.0 ** -20).astype(np.float32)
    s, c = O.sincos2pi(u)
    a = 2 * np.pi * u.astype(np.float64)
    assert np.abs(s - np.sin(a)).max() < 2e-7
    assert np.abs(c - np.cos(a)).max() < 2e-7
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 5e-7
    q = np.float32([0.0, 0.25, 0.5, 0.75])
    s, c = O.sincos2pi(q)
    np.testing.assert_array_equal(s, [0, 1, 0, -1])
    np.testing.assert_array_equal(c, [1, 0, -1, 0])

