"""CPU: the host case layer (derived constants and cut-cell tables) against
closed forms, quadrature and brute force.  Runs without a GPU: only the
host-side dsmc_case_describe of libcudadsmc.so is called."""
import math

import numpy as np
import pytest
from scipy import integrate

from cudadsmc_amd import abi, cases
from cudadsmc_amd.abi import DSMCError
from cudadsmc_amd.run import describe

ARGON = cases.ARGON
K = cases.BOLTZ


def _poly_area(xy):
    x, y = xy[:, 0], xy[:, 1]
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def test_vhs_constant_matches_gamma_closed_form():
    """K1 = pi d^2 (4 k Tref / m)^(omega - 1/2) / Gamma(5/2 - omega) (Bird
    1994 eq. 4.63 with m_r = m/2).  The reference's GammaFun (Run.cpp:9-27)
    returns 0.6257 for Gamma(1.69) instead of 0.9068; this test fails for it."""
    cfg = cases.box_equilibrium()
    info, _ = describe(cfg["params"])
    d, m, w, Tr = ARGON["diam"], ARGON["mass"], ARGON["Vtl"], ARGON["Tref"]
    k1 = math.pi * d * d * (4 * K * Tr / m) ** (w - 0.5) / math.gamma(2.5 - w)
    assert info.K1 == pytest.approx(k1, rel=1e-6)
    assert info.e_exp == pytest.approx(1.0 - w, rel=1e-7)
    assert abs(math.gamma(2.5 - w) - 0.9068) < 1e-3


def test_vhs_mean_sigma_cr_matches_collision_frequency():
    """n <sigma c_r> over the equilibrium relative-speed distribution equals
    the VHS equilibrium collision frequency 4 d^2 n sqrt(pi k Tref / m)
    (T/Tref)^(1-omega) (Bird 1994 eq. 4.64), by quadrature."""
    cfg = cases.box_equilibrium()
    info, _ = describe(cfg["params"])
    m, T, n = ARGON["mass"], 300.0, 1e20
    beta2 = m / (4 * K * T)                 # relative speed: reduced mass m/2
    f = lambda c: c * c * math.exp(-beta2 * c * c)
    s = lambda c: info.K1 * (c * c) ** info.e_exp * f(c)
    cmax = 12.0 / math.sqrt(beta2)
    mean = integrate.quad(s, 0, cmax, limit=200)[0] / integrate.quad(f, 0, cmax, limit=200)[0]
    assert n * mean == pytest.approx(cases.vhs_collision_frequency(n, T), rel=1e-5)
    assert cases.vhs_collision_frequency(n, T) == pytest.approx(29926.0, rel=2e-4)


def test_ntc_and_weight_constants():
    cfg = cases.box_equilibrium()
    p = cfg["params"]
    info, _ = describe(p)
    vcell = (p.L1 / p.N1) * (p.L2 / p.N2) * p.L3
    Fn = p.n_inf * vcell / p.ppc
    assert info.Fn == pytest.approx(Fn, rel=1e-12)
    assert info.ntc_coef == pytest.approx(Fn * p.tau / 2, rel=1e-6)
    vmp = math.sqrt(2 * K * p.T / ARGON["mass"])
    # reference initial (sigma c_r)max: Vmax pi d^2 with Vmax = 2 sqrt(8/pi) Vstd (Run.cpp:32-35, GPUCell.cu:224)
    assert info.sigcmax0 == pytest.approx(2 * math.sqrt(8 / math.pi) * vmp * math.pi * ARGON["diam"] ** 2, rel=1e-6)
    assert info.dt == p.tau
    assert info.ncell == p.N1 * p.N2


def test_inflow_rate_is_the_maxwellian_flux():
    """Expected inflow particles per step per face: the one-sided Maxwellian
    number flux (Bird 1994 eq. 4.22) times face area and dt over Fn, checked
    against direct quadrature of the flux integral."""
    cfg = cases.cylinder_mach10(nx=40, ny=20, n_points=32)
    p = cfg["params"]
    info, _ = describe(p, cfg["body"])
    m = ARGON["mass"]
    sig = math.sqrt(K * p.T / m)
    normals = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    lengths = [p.L2, p.L2, p.L1, p.L1]
    for f in range(4):
        u_n = p.V_jet * normals[f][0] + p.V_y * normals[f][1]
        g = lambda v: v * math.exp(-(v - u_n) ** 2 / (2 * sig * sig)) / math.sqrt(2 * math.pi) / sig
        flux = p.n_inf * integrate.quad(g, 0, max(u_n, 0) + 12 * sig, limit=400)[0]
        expect = flux * lengths[f] * p.L3 * p.tau / info.Fn
        assert info.inj_e[f] == pytest.approx(expect, rel=1e-7, abs=1e-12), f
        assert info.inj_a[f] == pytest.approx(u_n / math.sqrt(2 * K * p.T / m), rel=1e-6)
    assert info.inj_e[0] > 100 * info.inj_e[1]      # Mach 10: the upstream face dominates


def test_cut_volumes_sum_to_domain_minus_body():
    """Sum of fluid volumes = domain - polygon area (exact clipping), and
    -> domain - pi R^2 within the polygon's chord error."""
    for npts in (32, 256):
        cfg = cases.cylinder_mach10(nx=80, ny=40, n_points=npts)
        p, body = cfg["params"], cfg["body"]
        info, t = describe(p, body)
        total = p.L1 * p.L2 * p.L3
        assert t["vol"].sum() == pytest.approx(total - _poly_area(body) * p.L3, rel=1e-10)
        chord = math.pi * 0.1 ** 2 - _poly_area(body)
        assert 0 < chord < 0.5 * (2 * math.pi / npts) ** 2 * math.pi * 0.1 ** 2
        assert np.all((t["frac"] >= 0) & (t["frac"] <= 1))
        assert np.array_equal(t["cut"] == 1, (t["frac"] > 0) & (t["frac"] < 1))
        assert info.ncut == int(t["cut"].sum())


def test_wedge_cut_volume_is_exact():
    p = cases.make_params(L1=1.0, L2=0.5, N1=37, N2=23, n_inf=1e20, T=300.0, V_jet=1000.0,
                          face_bc=[abi.BC_STREAM] * 4, tau=1e-6, ppc=10.0)
    wedge = cases.wedge_polygon(0.3, 0.25, 0.4, 20.0)
    info, t = describe(p, wedge)
    h = 0.4 * math.tan(math.radians(20.0))
    assert _poly_area(wedge) == pytest.approx(0.4 * h, rel=1e-12)
    assert t["vol"].sum() == pytest.approx(0.5 - 0.4 * h, rel=1e-12)


def _point_in_poly(px, py, poly):
    inside = np.zeros(px.shape, bool)
    n = len(poly)
    for k in range(n):
        ax, ay = poly[k]
        bx, by = poly[(k + 1) % n]
        cross = (ay > py) != (by > py)
        xint = ax + (py - ay) * (bx - ax) / np.where(by != ay, by - ay, 1.0)
        inside ^= cross & (px < xint)
    return inside


def test_cut_volumes_match_monte_carlo_per_cell():
    """Per-cell fluid fraction of cut cells vs point sampling (brute force)."""
    cfg = cases.cylinder_mach10(nx=40, ny=20, n_points=48)
    p, body = cfg["params"], cfg["body"]
    info, t = describe(p, body)
    dx, dy = p.L1 / p.N1, p.L2 / p.N2
    rng = np.random.default_rng(5)
    cut = np.flatnonzero(t["cut"])
    assert cut.size > 10
    for c in cut:
        i, j = divmod(int(c), p.N2)                      # cell = j + i*N2 (CartesianMesh.cpp:171-181)
        u = rng.random((20000, 2))
        fx, fy = (i + u[:, 0]) * dx, (j + u[:, 1]) * dy
        frac = 1.0 - _point_in_poly(fx, fy, body).mean()
        assert abs(frac - t["frac"][c]) < 0.015, c
        fb = t["fbox"][4 * c:4 * c + 4]                 # fluid bounding box holds every fluid point
        out = ~_point_in_poly(fx, fy, body)
        assert fx[out].min() >= fb[0] - 1e-6 and fx[out].max() <= fb[2] + 1e-6
        assert fy[out].min() >= fb[1] - 1e-6 and fy[out].max() <= fb[3] + 1e-6
        # ... and every fluid point lies in a sub-box the fill mask marks
        sx = np.clip(((fx[out] - fb[0]) / (fb[2] - fb[0]) * 8).astype(int), 0, 7)
        sy = np.clip(((fy[out] - fb[1]) / (fb[3] - fb[1]) * 8).astype(int), 0, 7)
        bits = (np.uint64(t["fmask"][c]) >> (sx + 8 * sy).astype(np.uint64)) & np.uint64(1)
        assert np.all(bits == 1), c


def _seg_rect_dist(ax, ay, bx, by, x0, y0, x1, y1):
    """Distance from segment ab to the rectangle, by dense sampling of the segment."""
    s = np.linspace(0, 1, 2001)
    px, py = ax + s * (bx - ax), ay + s * (by - ay)
    ddx = np.maximum(np.maximum(x0 - px, 0), px - x1)
    ddy = np.maximum(np.maximum(y0 - py, 0), py - y1)
    return float(np.min(np.hypot(ddx, ddy)))


def test_segment_lists_cover_every_reachable_segment():
    """A cell's segment list holds every body segment within one margin of
    the cell (brute force), so a flight of at most one margin from inside the
    cell cannot cross an unlisted segment."""
    cfg = cases.cylinder_mach10(nx=40, ny=20, n_points=40)
    p, body = cfg["params"], cfg["body"]
    info, t = describe(p, body)
    dx, dy = p.L1 / p.N1, p.L2 / p.N2
    seg = t["seg"].reshape(-1, 6)
    cs, ci = t["cseg_start"], t["cseg_idx"]
    for c in range(info.ncell):
        i, j = divmod(c, p.N2)
        listed = set(ci[cs[c]:cs[c + 1]].tolist())
        for k, (ax, ay, bx, by, nx_, ny_) in enumerate(seg):
            d = _seg_rect_dist(ax, ay, bx, by, i * dx, j * dy, (i + 1) * dx, (j + 1) * dy)
            if d < info.margin * 0.999:
                assert k in listed, (c, k, d)
    # normals are unit and point out of the counter-clockwise body (into the fluid)
    cx, cy = 0.5 * p.L1, 0.5 * p.L2
    mid = np.stack([(seg[:, 0] + seg[:, 2]) / 2 - cx, (seg[:, 1] + seg[:, 3]) / 2 - cy], 1)
    assert np.allclose(np.hypot(seg[:, 4], seg[:, 5]), 1.0, atol=1e-6)
    assert np.all(np.sum(mid * seg[:, 4:6], 1) > 0)


@pytest.mark.parametrize("field,value,msg", [
    ("N1", 0, "N1, N2"),
    ("Vtl", 0.4, "omega"),
    ("tau", 0.0, "tau"),
    ("n_inf", -1.0, "n_inf"),
    ("mass", 0.0, "gas constants"),
])
def test_invalid_parameters_are_rejected(field, value, msg):
    cfg = cases.box_equilibrium()
    p = cfg["params"]
    setattr(p, field, value)
    with pytest.raises(DSMCError, match=msg):
        describe(p)


def test_bad_body_is_rejected():
    cfg = cases.cylinder_mach10(nx=40, ny=20, n_points=32)
    with pytest.raises(DSMCError, match="counter-clockwise"):
        describe(cfg["params"], cfg["body"][::-1].copy())
    with pytest.raises(DSMCError, match=">= 3"):
        describe(cfg["params"], cfg["body"][:2].copy())
    out = cfg["body"].copy()
    out[:, 1] += 0.12                  # the circle now crosses y = L2
    with pytest.raises(DSMCError, match="inside the domain"):
        describe(cfg["params"], out)


def test_sweep_wedges_stay_inside_the_domain():
    """Wedges of every half-angle of the sweep (up to 30 deg) keep their base
    at least 1 cm from the y faces."""
    for half in (10.0, 20.0, 26.0, 28.0, 30.0):
        c = cases.flow_case("wedge", nx=64, ny=32, half_angle_deg=half)
        b = c["body"]
        assert b[:, 1].min() >= 0.01 - 1e-12 and b[:, 1].max() <= 0.29 + 1e-12, half
        describe(c["params"], b)
    for c in cases.sweep(64):
        describe(c["params"], c["body"])


# ---------------------------------------------------------------- host-built tables
def _host_tables_rc(p, seg, frac=None, cut=None):
    import ctypes as C
    from cudadsmc_amd.abi import ptr
    lib = abi.load()
    seg = np.ascontiguousarray(seg, np.float64)
    t = abi.dsmc_cut_tables(seg.size // 6, ptr(seg, C.c_double), None, None)
    keep = [seg]
    if frac is not None:
        frac = np.ascontiguousarray(frac, np.float64); keep.append(frac); t.frac = ptr(frac, C.c_double)
    if cut is not None:
        cut = np.ascontiguousarray(cut, np.uint8); keep.append(cut); t.cut = ptr(cut, C.c_uint8)
    h = C.c_void_p()
    rc = lib.dsmc_create_from_tables(C.byref(p), C.byref(t), C.byref(h))
    if rc == abi.DSMC_OK:
        lib.dsmc_destroy(h)
    return rc, lib.dsmc_last_error()


def test_host_tables_are_checked_before_any_device():
    """dsmc_create_from_tables rejects an outline that does not close, a
    normal that points into the body, fractions outside [0, 1] and cut
    flags that disagree with the fractions -- on the host, with
    DSMC_ERR_GEOMETRY; valid tables get as far as the device check."""
    from cudadsmc_amd.run import outline_segments
    cfg = cases.cylinder_mach10(nx=40, ny=20, n_points=32)
    p, body = cfg["params"], cfg["body"]
    seg = outline_segments(body).reshape(-1, 6)
    info, t = describe(p, body)
    ok_rc, _ = _host_tables_rc(p, seg, t["frac"], t["cut"])
    assert ok_rc in (abi.DSMC_OK, abi.DSMC_ERR_NODEVICE)
    bad = seg.copy(); bad[3, 2] += 1e-3
    assert _host_tables_rc(p, bad) == (abi.DSMC_ERR_GEOMETRY, b"host tables: segment 3 does not end where the next one starts")
    bad = seg.copy(); bad[:, 4:] *= -1
    rc, msg = _host_tables_rc(p, bad)
    assert rc == abi.DSMC_ERR_GEOMETRY and b"unit normal" in msg
    f = t["frac"].copy(); f[5] = 1.5
    assert _host_tables_rc(p, seg, f)[0] == abi.DSMC_ERR_GEOMETRY
    c = t["cut"].copy(); c[np.argmax(c)] = 0
    rc, msg = _host_tables_rc(p, seg, t["frac"], c)
    assert rc == abi.DSMC_ERR_GEOMETRY and b"cut flags" in msg
