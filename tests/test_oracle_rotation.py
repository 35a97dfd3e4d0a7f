"""Pins of the oracle's rotational-energy restatement (zeta = 2, the
reference's N2: src/Param.h:14; the Larsen-Borgnakke scheme written out and
commented in src/cell/GPUCell.cu:256-275,316-328; Random::Rotationalenergy-
Sample, src/random/Random.cpp:48-51).

Each check is fixed by mathematics the scheme must obey, not by its own
formula:
  - every collision conserves momentum and the total energy
    (m/2)(v1^2 + v2^2) + E_r1 + E_r2 (the reference's own energy bookkeeping,
    with the full mass in E_t, fails this);
  - the translational share x = E_t'/E_c of an exchanging collision follows
    Beta(5/2 - omega, zeta) and the rotational split E_r1'/E_r' follows
    Beta(1, 1) (Kolmogorov-Smirnov against scipy's distributions);
  - the fraction of exchanging collisions is rot_prob (binomial);
  - in a closed box T_tr and T_rot relax towards the energy-conserving
    common temperature (3 T_tr0 + zeta T_rot0)/(3 + zeta) at the rate the
    Larsen-Borgnakke energy balance gives per collision,
    d(T_tr - T_rot)/dc = -(1 + zeta/3) p a/(a + zeta) (T_tr - T_rot),
    a = 5/2 - omega, c = collisions per molecule, p = rot_prob;
  - at equilibrium E_rot is exponential with mean kT (two degrees of freedom);
  - diffuse walls, the inflow and the fill emit E_rot ~ Exp(kT) at their own
    temperature.
"""
import math

import numpy as np
import pytest
from scipy import stats

from cudadsmc_amd import abi, cases
from oracle import oracle as O

N2 = cases.NITROGEN


def n2_params(**kw):
    base = dict(mass=N2["mass"], diam=N2["diam"], Vtl=N2["Vtl"], Tref=N2["Tref"], zeta=2, rot_prob=1.0)
    base.update(kw)
    return cases.make_params(**base)


def _pair_cells(rot_prob, K=40000, T=300.0, seed=7):
    """K cells of two particles each, one NTC candidate per cell (the
    remainder carry set so that M = 1), Maxwellian velocities and
    exponential rotational energies at T.  Returns the state before and after
    one collide() and the accepted-pair mask."""
    p = n2_params(L1=1.0, L2=1e-3, N1=K, N2=1, n_inf=1e20, T=T, V_jet=0.0, wall_bc=abi.BC_SPECULAR,
                  face_bc=[abi.BC_SPECULAR] * 4, tau=4e-5, ppc=2.0, rot_prob=rot_prob)
    oc, info, tables = O.case_from_params(p)
    rng = np.random.default_rng(seed)
    n = 2 * K
    kt_m = p.boltz * T / p.mass
    sig = math.sqrt(kt_m)
    v = (rng.standard_normal((3, n)) * sig).astype(np.float32)
    er = (rng.exponential(kt_m, n)).astype(np.float32)
    start = np.arange(0, n + 1, 2, dtype=np.int32)
    perm = np.arange(n, dtype=np.int32)
    vol = tables["vol_ntc"].astype(np.float64)
    smax = (0.5 * vol / (2.0 * info.ntc_coef)).astype(np.float32)
    mrem = np.full(K, 0.6, np.float32)
    before = dict(vx=v[0].copy(), vy=v[1].copy(), vz=v[2].copy(), er=er.copy())
    vx, vy, vz, e = v[0].copy(), v[1].copy(), v[2].copy(), er.copy()
    ncoll = oc.collide(0, start, perm, vx, vy, vz, smax, mrem, er=e)
    after = dict(vx=vx, vy=vy, vz=vz, er=e)
    assert ncoll.max() <= 1
    acc = ncoll.astype(bool)
    return p, before, after, acc


def _pairs(d, key):
    a = d[key].astype(np.float64)
    return a[0::2], a[1::2]


@pytest.fixture(scope="module")
def every_collision():
    return _pair_cells(1.0)


def test_collisions_conserve_momentum_and_energy(every_collision):
    p, b, a, acc = every_collision
    assert acc.sum() > 10000
    for k in ("vx", "vy", "vz"):
        s0 = sum(_pairs(b, k))
        s1 = sum(_pairs(a, k))
        np.testing.assert_allclose(s1, s0, atol=2e-4 * math.sqrt(p.boltz * p.T / p.mass))

    def energy(d):
        e = 0.0
        for k in ("vx", "vy", "vz"):
            u, w = _pairs(d, k)
            e = e + 0.5 * (u * u + w * w)
        e1, e2 = _pairs(d, "er")
        return e + e1 + e2
    E0, E1 = energy(b), energy(a)
    rel = np.abs(E1 - E0) / E0
    assert rel.max() < 2e-6, rel.max()
    # the reference's bookkeeping, E_t = m g^2 / 2 with the full mass, would
    # change the pair energy by a fraction of order one
    assert np.all(np.isfinite(E1)) and np.all(a["er"] >= 0.0)


def _shares(b, a, acc):
    g2b = sum((_pairs(b, k)[0] - _pairs(b, k)[1]) ** 2 for k in ("vx", "vy", "vz"))
    g2a = sum((_pairs(a, k)[0] - _pairs(a, k)[1]) ** 2 for k in ("vx", "vy", "vz"))
    e1b, e2b = _pairs(b, "er")
    e1a, e2a = _pairs(a, "er")
    ec = 0.25 * g2b + e1b + e2b
    exch = acc & (np.abs((e1a + e2a) - (e1b + e2b)) > 1e-6 * ec)
    return g2b, g2a, ec, e1a, e2a, exch


def test_translational_share_is_beta(every_collision):
    p, b, a, acc = every_collision
    g2b, g2a, ec, e1a, e2a, exch = _shares(b, a, acc)
    # every accepted collision exchanges with rot_prob = 1
    assert exch.sum() >= acc.sum() - 5
    x = (0.25 * g2a / ec)[exch]
    alpha = 2.5 - p.Vtl
    ks = stats.kstest(x, stats.beta(alpha, 2.0).cdf)
    assert ks.pvalue > 1e-4, ks
    # a plausible mistake, one factor of the product dropped, draws Beta(a, 1)
    assert stats.kstest(x, stats.beta(alpha, 1.0).cdf).pvalue < 1e-10


def test_rotational_split_is_uniform(every_collision):
    p, b, a, acc = every_collision
    g2b, g2a, ec, e1a, e2a, exch = _shares(b, a, acc)
    tot = e1a + e2a
    ok = exch & (tot > 0)
    f = (e1a / tot)[ok]
    ks = stats.kstest(f, stats.uniform().cdf)
    assert ks.pvalue > 1e-4, ks
    # the reference's live code halves the energy between the two
    assert np.mean(np.abs(f - 0.5) < 1e-6) < 0.01


@pytest.mark.parametrize("rot_prob", [0.3, 0.05])
def test_exchange_fraction_is_rot_prob(rot_prob):
    p, b, a, acc = _pair_cells(rot_prob, K=30000, seed=11)
    g2b, g2a, ec, e1a, e2a, exch = _shares(b, a, acc)
    na = int(acc.sum())
    k = int(exch.sum())
    sd = math.sqrt(na * rot_prob * (1 - rot_prob))
    assert abs(k - na * rot_prob) < 5 * sd, (k, na)
    # elastic ones keep |g| and both energies
    el = acc & ~exch
    np.testing.assert_allclose(np.sqrt(g2a[el]), np.sqrt(g2b[el]), rtol=2e-6)
    e1b, e2b = _pairs(b, "er")
    np.testing.assert_array_equal(e1a[el], e1b[el])
    np.testing.assert_array_equal(e2a[el], e2b[el])


def _box(rot_prob, T0=400.0, rot_scale=0.25, ppc=50, n_cells=32, seed=3):
    """Closed specular N2 box filled at T0, rotational energies scaled by
    rot_scale (T_rot0 = rot_scale T0)."""
    n = 1e20
    nu = cases.vhs_collision_frequency(n, T0, gas=N2)
    p = n2_params(L1=1e-3, L2=1e-3, N1=n_cells, N2=n_cells, n_inf=n, T=T0, V_jet=0.0,
                  wall_bc=abi.BC_SPECULAR, face_bc=[abi.BC_SPECULAR] * 4, tau=0.2 / nu, ppc=float(ppc),
                  seed=seed, rot_prob=rot_prob)
    oc, info, tables = O.case_from_params(p)
    C = info.ncell
    expected = np.full(C, info.ppc_eff) * tables["frac"]
    sig = np.full(C, info.inj_sig, np.float32)
    parts, _, err = oc.init_fill(expected, tables["cut"], np.zeros(C, np.float32), np.zeros(C, np.float32), sig,
                                 int(expected.sum() * 1.2) + 1000)
    assert err == 0
    parts["er"] = (parts["er"] * np.float32(rot_scale)).astype(np.float32)
    st = O.OracleState(oc, parts, parts["x"].shape[0] + 16, info.sigcmax0)
    return p, st


def _temps(p, parts):
    v = np.stack([parts[k].astype(np.float64) for k in ("vx", "vy", "vz")])
    u = v.mean(axis=1, keepdims=True)
    Ttr = p.mass / (3 * p.boltz) * np.mean(np.sum((v - u) ** 2, axis=0))
    Trot = p.mass / p.boltz * np.mean(parts["er"].astype(np.float64))
    return Ttr, Trot


def test_box_relaxes_at_the_larsen_borgnakke_rate():
    rp = 0.2
    p, st = _box(rp)
    parts = st.particles()
    N = parts["x"].shape[0]
    Ttr0, Trot0 = _temps(p, parts)
    Teq = (3 * Ttr0 + 2 * Trot0) / 5
    c, dT = [0.0], [Ttr0 - Trot0]
    coll0 = 0
    for s in range(120):
        st.step(sample=False, nthreads=4)
        Ttr, Trot = _temps(p, st.particles())
        c.append(2.0 * st.st.total_coll / N)
        dT.append(Ttr - Trot)
    c, dT = np.array(c), np.array(dT)
    # energy: 3/2 T_tr + T_rot per molecule is conserved by every collision
    Ttr, Trot = _temps(p, st.particles())
    assert abs((3 * Ttr + 2 * Trot) / 5 / Teq - 1) < 2e-4
    # rate: log dT linear in collisions per molecule, slope -(5/3) p a/(a+2)
    a = 2.5 - p.Vtl
    k_th = (1 + 2 / 3) * rp * a / (a + 2)
    use = dT > 0.1 * dT[0]
    slope = np.polyfit(c[use], np.log(dT[use]), 1)[0]
    assert abs(-slope / k_th - 1) < 0.06, (slope, k_th)


def test_equilibrium_rotational_energy_is_exponential():
    p, st = _box(1.0, rot_scale=1.0, ppc=40, n_cells=24, seed=5)
    for s in range(200):
        st.step(sample=False, nthreads=4)
    parts = st.particles()
    Ttr, Trot = _temps(p, parts)
    assert abs(Trot / Ttr - 1) < 0.01, (Ttr, Trot)
    x = parts["er"].astype(np.float64) * p.mass / (p.boltz * Ttr)
    assert stats.kstest(x, stats.expon().cdf).pvalue > 1e-4
    for k in ("vx", "vy", "vz"):
        z = parts[k].astype(np.float64) / math.sqrt(p.boltz * Ttr / p.mass)
        assert stats.kstest(z, stats.norm().cdf).pvalue > 1e-4


def test_diffuse_wall_emits_wall_rotational_energy():
    Tw = 900.0
    p = n2_params(L1=1.0, L2=1.0, N1=4, N2=4, n_inf=1e20, T=300.0, V_jet=0.0, wall_bc=abi.BC_SPECULAR,
                  face_bc=[abi.BC_DIFFUSE, abi.BC_SPECULAR, abi.BC_SPECULAR, abi.BC_SPECULAR],
                  face_T=[Tw, 0.0, 0.0, 0.0], tau=6e-4, ppc=2.0)
    oc, info, _ = O.case_from_params(p)
    n = 60000
    rng = np.random.default_rng(2)
    # every particle reaches x = 0 within the step and leaves it once
    s = dict(x=rng.uniform(0.01, 0.2, n).astype(np.float32), y=rng.uniform(0.1, 0.9, n).astype(np.float32),
             vx=np.full(n, -400.0, np.float32), vy=np.zeros(n, np.float32), vz=np.zeros(n, np.float32),
             er=np.full(n, 123.0, np.float32), pid=np.arange(n, dtype=np.uint32))
    dead, err = oc.move(s, 5)
    assert err == 0 and not dead.any()
    hit = s["vx"] > 0
    assert hit.mean() > 0.99
    x = s["er"][hit].astype(np.float64) / (p.boltz * Tw / p.mass)
    assert stats.kstest(x, stats.expon().cdf).pvalue > 1e-4
    # specular faces keep it
    q = dict(x=np.full(1000, 0.95, np.float32), y=np.full(1000, 0.5, np.float32),
             vx=np.full(1000, 400.0, np.float32), vy=np.zeros(1000, np.float32), vz=np.zeros(1000, np.float32),
             er=np.full(1000, 7.0, np.float32), pid=np.arange(1000, dtype=np.uint32))
    oc.move(q, 5)
    assert np.all(q["vx"] < 0) and np.all(q["er"] == np.float32(7.0))


def test_inflow_and_fill_draw_freestream_rotational_energy():
    p = n2_params(L1=0.2, L2=0.1, N1=20, N2=10, n_inf=1e21, T=250.0, V_jet=800.0, tau=2e-5, ppc=400.0)
    oc, info, tables = O.case_from_params(p)
    kt_m = p.boltz * p.T / p.mass
    J = oc.inject_counts(3)
    out, dead, err = oc.inject(3, 1000, int(J.sum()))
    assert err == 0 and J.sum() > 5000
    x = out["er"].astype(np.float64) / kt_m
    assert stats.kstest(x, stats.expon().cdf).pvalue > 1e-4
    C = info.ncell
    expected = np.full(C, info.ppc_eff) * tables["frac"]
    parts, _, err = oc.init_fill(expected, tables["cut"], np.zeros(C, np.float32), np.zeros(C, np.float32),
                                 np.full(C, info.inj_sig, np.float32), int(expected.sum() * 1.2) + 100)
    x = parts["er"].astype(np.float64) / (float(np.float32(info.inj_sig)) ** 2)
    assert stats.kstest(x, stats.expon().cdf).pvalue > 1e-4


def test_monatomic_case_has_no_rotational_state():
    p = cases.make_params(L1=1e-3, L2=1e-3, N1=4, N2=4, n_inf=1e20, T=300.0, V_jet=0.0, tau=1e-7, ppc=5.0)
    oc, info, _ = O.case_from_params(p)
    assert info.zeta == 0 and info.sample_blocks == 6
    p2 = n2_params(L1=1e-3, L2=1e-3, N1=4, N2=4, n_inf=1e20, T=300.0, V_jet=0.0, tau=1e-7, ppc=5.0)
    _, info2, _ = O.case_from_params(p2)
    assert info2.zeta == 2 and info2.sample_blocks == 7
    assert info2.lb_ia == pytest.approx(1 / (2.5 - N2["Vtl"]), rel=1e-6)
    assert info2.lb_ib == pytest.approx(1 / (3.5 - N2["Vtl"]), rel=1e-6)
    assert info2.inj_ekt == pytest.approx(p2.boltz * 300.0 / p2.mass, rel=1e-6)
