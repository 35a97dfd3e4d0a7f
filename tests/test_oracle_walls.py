"""CPU: pins of the oracle's wall and inflow physics against an independent
float64 brute force and kinetic-theory moments.

  - inflow on each of the four stream faces (GPUParticle.cu:375-422, restated
    as the Maxwellian flux through every stream face): entry position on the
    face, inward normal speed with the flux-weighted law of the drift along
    that face's inward normal, the tangential drift (V_y on x-faces, V_jet on
    y-faces), Gaussian tangential and z components, independent;
  - body walls (GPUSegment.cuh:20-95, Segment.cpp:18-53): a ray-segment brute
    force over every segment of the body (no cell lists) gives the earliest
    hit of each flight, including flights several cells long and concave
    bodies where one flight crosses several wall segments; the specular law
    angle in = angle out (Segment.cpp:40); diffuse re-emission moments off a
    slanted body segment (WallBoundary.cpp:33-43 with fix D7).

The brute force is written from the geometry, not from the oracle's code:
float64, every segment of the body tested for every event, events ordered
by time.  Particles whose outcome FP32 rounding could flip (two events
within 1e-6 of the flight time, a hit within 1e-4 of a segment end, a
grazing hit, a start on a wall line) are left out of the comparison."""
import math

import numpy as np
import pytest
from scipy import integrate

from cudadsmc_amd import abi, cases
from oracle import oracle as O

K = cases.BOLTZ
M = cases.ARGON["mass"]


def _z(x, sigma):
    return abs(x) / sigma


# ---------------------------------------------------------------- inflow, every face
def _inflow_case(V_jet, V_y, T=300.0, n_cells=32, L=1.0, tau=2e-6, ppc=2000.0, seed=31):
    p = cases.make_params(L1=L, L2=L, N1=n_cells, N2=n_cells, n_inf=1e20, T=T, V_jet=V_jet, V_y=V_y,
                          wall_bc=abi.BC_SPECULAR, face_bc=[abi.BC_STREAM] * 4, tau=tau, ppc=ppc, seed=seed)
    oc, info, _ = O.case_from_params(p, None)
    return oc, info, p


def _collect_inflow(oc, per_face_min):
    """Inflow particles of successive steps, split by face (the oracle creates
    a step's particles face by face, J[0] of face 0 first)."""
    out = {f: {k: [] for k in ("x", "y", "vx", "vy", "vz", "dead")} for f in range(4)}
    nxt, s = 0, 0
    while min(sum(len(a) for a in out[f]["x"]) for f in range(4)) < per_face_min:
        J = oc.inject_counts(s)
        parts, dead, err = oc.inject(s, nxt, int(J.sum()))
        assert err == 0
        o = 0
        for f in range(4):
            sl = slice(o, o + int(J[f]))
            for k in ("x", "y", "vx", "vy", "vz"):
                out[f][k].append(parts[k][sl].astype(np.float64))
            out[f]["dead"].append(dead[sl])
            o += int(J[f])
        nxt += int(J.sum())
        s += 1
        assert s < 5000
    return {f: {k: np.concatenate(v) for k, v in d.items()} for f, d in out.items()}


def _flux_moments(a_drift, sig):
    """Mean and s.d. of the inward normal speed of molecules crossing a face,
    density v f(v) on v > 0 with f the Maxwellian of mean a_drift."""
    w = lambda v: v * math.exp(-(v - a_drift) ** 2 / (2 * sig * sig))
    hi = max(a_drift, 0.0) + 12 * sig
    z = integrate.quad(w, 0, hi, limit=200)[0]
    m1 = integrate.quad(lambda v: v * w(v), 0, hi, limit=200)[0] / z
    m2 = integrate.quad(lambda v: v * v * w(v), 0, hi, limit=200)[0] / z
    return m1, math.sqrt(m2 - m1 * m1)


@pytest.mark.parametrize("V_jet,V_y", [(300.0, -200.0), (-150.0, 250.0)])
def test_inflow_on_every_face(V_jet, V_y):
    """Each stream face f injects at its own boundary, inward, with the
    flux-weighted normal speed of the drift component V.n_f, the tangential
    drift of the other velocity component, thermal tangential and z parts,
    uniform entry times and positions along the face."""
    oc, info, p = _inflow_case(V_jet, V_y)
    sig = math.sqrt(K * p.T / M)
    # inflow rate of each face: one-sided Maxwellian flux (Bird 1994 eq. 4.22)
    # by quadrature, with the drift of that face's inward normal
    normals = [(1.0, 0.0), (-1.0, 0.0), (0.0, 1.0), (0.0, -1.0)]
    lengths = [p.L2, p.L2, p.L1, p.L1]
    for f in range(4):
        un = V_jet * normals[f][0] + V_y * normals[f][1]
        g = lambda v: v * math.exp(-(v - un) ** 2 / (2 * sig * sig)) / math.sqrt(2 * math.pi) / sig
        flux = p.n_inf * integrate.quad(g, 0, max(un, 0) + 12 * sig, limit=400)[0]
        assert info.inj_e[f] == pytest.approx(flux * lengths[f] * p.L3 * p.tau / info.Fn, rel=1e-7), f
    d = _collect_inflow(oc, 40000)
    dt = np.float64(np.float32(info.dt))
    L1, L2 = float(np.float32(p.L1)), float(np.float32(p.L2))
    for f in range(4):
        q = d[f]
        live = q["dead"] == 0
        assert live.mean() > 0.995, f"face {f}: {1 - live.mean():.4f} of its inflow leaves at once"
        x, y, vx, vy, vz = (q[k][live] for k in ("x", "y", "vx", "vy", "vz"))
        n = x.size
        if f < 2:
            vn, vt, tpos, L, udrift, tdrift = (vx if f == 0 else -vx), vy, y, L2, (V_jet if f == 0 else -V_jet), V_y
            depth = x if f == 0 else L1 - x
        else:
            vn, vt, tpos, L, udrift, tdrift = (vy if f == 2 else -vy), vx, x, L1, (V_y if f == 2 else -V_y), V_jet
            depth = y if f == 2 else L2 - y
        assert np.all(vn > 0), f"face {f}: a particle enters outward"
        # entry time: depth / v_n is the flight time, uniform on (0, dt)
        frac = depth / (vn * dt)
        tol = 1e-4 + 2e-7 * L / (vn * dt)          # FP32 spacing of positions near x = L
        assert np.all((frac >= -tol) & (frac <= 1 + tol)), f"face {f}: entry off the face"
        assert _z(frac.mean() - 0.5, math.sqrt(1 / 12 / n)) < 4, f
        assert np.all((tpos >= 0) & (tpos <= L)), f
        assert _z(tpos.mean() - L / 2, L * math.sqrt(1 / 12 / n)) < 4, f
        m1, sd = _flux_moments(udrift, sig)
        assert _z(vn.mean() - m1, sd / math.sqrt(n)) < 4, (f, vn.mean(), m1)
        assert _z(vn.std() - sd, sd / math.sqrt(2 * n)) < 4, f
        assert _z(vt.mean() - tdrift, sig / math.sqrt(n)) < 4, (f, vt.mean(), tdrift)
        assert _z(vt.var() - sig * sig, sig * sig * math.sqrt(2 / n)) < 4, f
        assert _z(vz.mean(), sig / math.sqrt(n)) < 4, f
        assert _z(np.mean(vz * vz) - sig * sig, sig * sig * math.sqrt(2 / n)) < 4, f
        for a, b in ((vt, vz), (vn, vt), (vn, vz)):
            assert abs(np.corrcoef(a, b)[0, 1]) < 4 / math.sqrt(n), f


# ---------------------------------------------------------------- body walls: brute force
def _star(cx, cy, r_out, r_in, k=5, phase=0.1):
    a = phase + np.pi * np.arange(2 * k) / k
    r = np.where(np.arange(2 * k) % 2 == 0, r_out, r_in)
    return np.stack([cx + r * np.cos(a), cy + r * np.sin(a)], axis=1)


def _inside(px, py, poly):
    inside = np.zeros(px.shape, bool)
    for k in range(len(poly)):
        ax, ay = poly[k]
        bx, by = poly[(k + 1) % len(poly)]
        cross = (ay > py) != (by > py)
        xint = ax + (py - ay) * (bx - ax) / np.where(by != ay, by - ay, 1.0)
        inside ^= cross & (px < xint)
    return inside


def _seg_dist(px, py, seg):
    """Distance of every point to every segment (n, nseg)."""
    ax, ay, bx, by = (seg[:, k][None, :] for k in range(4))
    ex, ey = bx - ax, by - ay
    t = np.clip(((px[:, None] - ax) * ex + (py[:, None] - ay) * ey) / (ex * ex + ey * ey), 0, 1)
    return np.hypot(px[:, None] - ax - t * ex, py[:, None] - ay - t * ey)


def brute_move(x, y, vx, vy, t, seg, L1, L2, face_bc, body_specular=True, max_events=64):
    """Flight of duration t with specular body segments and specular or
    stream faces, every event found over all segments and faces, earliest
    first (float64).  Returns final x, y, vx, vy, dead, a mask of particles
    whose outcome sits within rounding of a branch, and per-particle event
    records (list of (kind, segment) tuples)."""
    x, y, vx, vy = (np.array(a, np.float64) for a in (x, y, vx, vy))
    t = np.full(x.shape, float(t))
    ax, ay, bx, by, nx, ny = (seg[:, k][None, :].astype(np.float64) for k in range(6))
    ex, ey = bx - ax, by - ay
    n = x.size
    dead = np.zeros(n, bool)
    amb = np.zeros(n, bool)
    done = np.zeros(n, bool)
    t0 = t.copy()
    events = [[] for _ in range(n)]
    for _ in range(max_events):
        act = ~done
        if not act.any():
            break
        # body segments
        vn = vx[:, None] * nx + vy[:, None] * ny
        den = vx[:, None] * ey - vy[:, None] * ex
        wx, wy = ax - x[:, None], ay - y[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            s = (wx * ey - wy * ex) / den
            u = (wx * vy[:, None] - wy * vx[:, None]) / den
        ok = (vn < 0) & (den != 0) & (s > 0) & (s < t[:, None]) & (u >= 0) & (u <= 1)
        near_end = (vn < 0) & (den != 0) & (s > -1e-9 * t0[:, None]) & (s < t[:, None] * (1 + 1e-6)) & \
            ((np.abs(u) < 1e-4) | (np.abs(u - 1) < 1e-4))
        sb = np.where(ok, s, np.inf)
        # faces
        with np.errstate(divide="ignore", invalid="ignore"):
            fs = np.stack([np.where(vx < 0, (0 - x) / vx, np.inf), np.where(vx > 0, (L1 - x) / vx, np.inf),
                           np.where(vy < 0, (0 - y) / vy, np.inf), np.where(vy > 0, (L2 - y) / vy, np.inf)], 1)
        fs = np.where(fs < t[:, None], fs, np.inf)
        allt = np.concatenate([fs, sb], 1)
        kind = np.argmin(allt, 1)
        best = allt[np.arange(n), kind]
        srt = np.sort(allt, 1)
        hit = act & np.isfinite(best)
        # rounding-sensitive: two events close in time, a vertex hit, grazing
        # incidence, an event at the very end of the flight
        with np.errstate(invalid="ignore"):
            amb |= act & np.isfinite(srt[:, 1]) & (srt[:, 1] - srt[:, 0] < 1e-6 * t0)
        amb |= act & near_end.any(1)
        amb |= act & (np.abs(best - t) < 1e-6 * t0)
        k = np.clip(kind - 4, 0, seg.shape[0] - 1)
        spd = np.hypot(vx, vy)
        vnk = vn[np.arange(n), k]
        amb |= hit & (kind >= 4) & (np.abs(vnk) < 1e-3 * spd)
        # no event: free flight to the end
        fin = act & ~hit
        x[fin] += vx[fin] * t[fin]
        y[fin] += vy[fin] * t[fin]
        done |= fin
        h = np.flatnonzero(hit)
        b = best[h]
        x[h] += vx[h] * b
        y[h] += vy[h] * b
        t[h] -= b
        for i, kk in zip(h, kind[h]):
            events[i].append((int(kk), int(kk) - 4 if kk >= 4 else -1))
        kk = kind[h]
        for f in range(4):
            sel = h[kk == f]
            if f == 0: x[sel] = 0.0
            elif f == 1: x[sel] = L1
            elif f == 2: y[sel] = 0.0
            else: y[sel] = L2
            if face_bc[f] == abi.BC_STREAM:
                dead[sel] = True
                done[sel] = True
            elif f < 2:
                vx[sel] = -vx[sel]
            else:
                vy[sel] = -vy[sel]
        sel = h[kk >= 4]
        ks = kk[kk >= 4] - 4
        if body_specular:
            v_n = vx[sel] * seg[ks, 4] + vy[sel] * seg[ks, 5]
            vx[sel] = vx[sel] - 2 * v_n * seg[ks, 4]
            vy[sel] = vy[sel] - 2 * v_n * seg[ks, 5]
        else:
            done[sel] = True      # caller reads the first hit only
    amb |= ~done
    return x, y, vx, vy, dead, amb, events


def _specular_body_case(body, L, n_cells, tau, faces=abi.BC_SPECULAR, seed=41):
    p = cases.make_params(L1=L, L2=L, N1=n_cells, N2=n_cells, n_inf=1e20, T=300.0, V_jet=0.0,
                          wall_bc=abi.BC_SPECULAR, face_bc=[faces] * 4, tau=tau, ppc=10.0, seed=seed)
    oc, info, tables = O.case_from_params(p, body)
    return oc, info, tables, p


def _flights(rng, n, body, L, lmin, lmax, tau, seg):
    """Particles outside the body (and off its wall lines), in a band of
    width lmax around the body's bounding box, with flights of length
    U[lmin, lmax] in random directions."""
    lo = np.maximum(body.min(0) - lmax, 0.0)
    hi = np.minimum(body.max(0) + lmax, L)
    xs, ys = [], []
    while sum(len(a) for a in xs) < n:
        px = lo[0] + rng.random(4 * n) * (hi[0] - lo[0])
        py = lo[1] + rng.random(4 * n) * (hi[1] - lo[1])
        ok = ~_inside(px, py, body)
        ok &= _seg_dist(px, py, seg).min(1) > 1e-6 * L
        xs.append(px[ok]); ys.append(py[ok])
    x = np.concatenate(xs)[:n].astype(np.float32)
    y = np.concatenate(ys)[:n].astype(np.float32)
    th = rng.random(n) * 2 * np.pi
    spd = (lmin + (lmax - lmin) * rng.random(n)) / tau
    vx = (spd * np.cos(th)).astype(np.float32)
    vy = (spd * np.sin(th)).astype(np.float32)
    vz = rng.normal(0, 300, n).astype(np.float32)
    return dict(x=x, y=y, vx=vx, vy=vy, vz=vz, pid=np.arange(n, dtype=np.uint32))


def _compare_with_brute_force(oc, info, tables, p, parts, what):
    seg = tables["seg"].reshape(-1, 6).astype(np.float64)
    s = {k: v.copy() for k, v in parts.items()}
    dead, err = oc.move(s, 5)
    assert err == 0, what
    f32 = lambda v: float(np.float32(v))
    bx, by, bvx, bvy, bdead, amb, ev = brute_move(parts["x"], parts["y"], parts["vx"], parts["vy"],
                                                  f32(info.dt), seg, f32(p.L1), f32(p.L2), list(p.face_bc))
    keep = ~amb
    assert keep.mean() > 0.98, f"{what}: {(~keep).sum()} particles near a rounding branch"
    np.testing.assert_array_equal(dead[keep].astype(bool), bdead[keep], err_msg=what)
    L = max(p.L1, p.L2)
    for a, b, tol, name in ((s["x"], bx, 2e-5 * L, "x"), (s["y"], by, 2e-5 * L, "y")):
        err_ = np.abs(a[keep].astype(np.float64) - b[keep])
        assert err_.max() < tol, f"{what}: {name} off by {err_.max():.3g} ({(err_ >= tol).sum()} particles)"
    spd = np.hypot(parts["vx"], parts["vy"]).astype(np.float64)
    for a, b, name in ((s["vx"], bvx, "vx"), (s["vy"], bvy, "vy")):
        err_ = np.abs(a[keep].astype(np.float64) - b[keep]) / spd[keep]
        assert err_.max() < 1e-4, f"{what}: {name} off by {err_.max():.3g} of the speed"
    np.testing.assert_array_equal(s["vz"], parts["vz"])      # specular walls keep vz
    return s, ev, keep, seg


def _comb(x0=0.2, x1=0.8, y0=0.2, yb=0.3, y1=0.8, teeth=8, width=0.02):
    """Counter-clockwise comb: a base [x0, x1] x [y0, yb] carrying thin
    teeth of the given width up to y1 (concave; a flight across the teeth
    crosses several front-facing walls)."""
    pitch = (x1 - x0 - width) / (teeth - 1)
    pts = [(x0, y0), (x1, y0), (x1, yb)]
    for k in reversed(range(teeth)):
        a = x0 + k * pitch
        b = a + width
        if k == teeth - 1 and abs(b - x1) < 1e-12:
            pts[-1] = (x1, y1)
            pts.append((a, y1))
            pts.append((a, yb))
            continue
        pts += [(b, yb), (b, y1), (a, y1)]
        if k > 0:
            pts.append((a, yb))
    return np.array(pts, np.float64)


@pytest.mark.parametrize("shape", ["comb", "star"])
def test_earliest_wall_hit_against_brute_force(shape):
    """A concave body (a comb of thin teeth; a five-point star) in a 32x32
    box, specular everywhere, flights 0.5 to 6 cells long: final position,
    velocity and deletion of every particle equal a brute force over all
    segments.  The sample is checked to hold flights that cross several
    front-facing walls (so 'the earliest hit' matters), particles with
    several wall events in one step, and hits on segments beyond the cell
    list margin of the flight's start."""
    body = _comb() if shape == "comb" else _star(0.5, 0.5, 0.32, 0.12)
    oc, info, tables, p = _specular_body_case(body, 1.0, 32, 1e-4)
    rng = np.random.default_rng(7)
    parts = _flights(rng, 60000, body, 1.0, 0.5 * info.margin, 6 * info.margin, p.tau, tables["seg"].reshape(-1, 6))
    s, ev, keep, seg = _compare_with_brute_force(oc, info, tables, p, parts, shape)
    # coverage of the sample
    body_hits = np.array([sum(1 for k, _ in e if k >= 4) for e in ev])
    x, y, vx, vy = (parts[k].astype(np.float64) for k in ("x", "y", "vx", "vy"))
    t = np.float64(np.float32(info.dt))
    ex, ey = seg[:, 2] - seg[:, 0], seg[:, 3] - seg[:, 1]
    vn = vx[:, None] * seg[:, 4] + vy[:, None] * seg[:, 5]
    den = vx[:, None] * ey - vy[:, None] * ex
    wx, wy = seg[:, 0] - x[:, None], seg[:, 1] - y[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        sh = (wx * ey - wy * ex) / den
        uh = (wx * vy[:, None] - wy * vx[:, None]) / den
    crossing = ((vn < 0) & (sh > 0) & (sh < t) & (uh >= 0) & (uh <= 1)).sum(1)
    # hits on a segment absent from the start cell's list (reach > margin)
    i = np.minimum((parts["x"] / np.float32(info.dx)).astype(int), p.N1 - 1)
    j = np.minimum((parts["y"] / np.float32(info.dy)).astype(int), p.N2 - 1)
    c = j + i * p.N2
    cs, ci = tables["cseg_start"], tables["cseg_idx"]
    far = 0
    for q in np.flatnonzero(keep & (body_hits >= 1)):
        k0 = next(k for kind, k in ev[q] if kind >= 4)
        if k0 not in ci[cs[c[q]]:cs[c[q] + 1]]:
            far += 1
    need = dict(comb=(8000, 1000, 1500, 2000), star=(3000, 150, 0, 1000))[shape]
    got = ((body_hits[keep] >= 1).sum(), (body_hits[keep] >= 2).sum(), (crossing[keep] >= 2).sum(), far)
    assert all(g >= n_ for g, n_ in zip(got, need)), (got, need)


@pytest.mark.parametrize("geometry", ["cylinder", "wedge"])
def test_specular_body_reflection(geometry):
    """Cylinder (128-gon) and wedge with specular walls, the product's own
    cut-cell tables, flights up to 6 cells: every particle's outcome equals
    the brute force, and on every single-hit flight the wall law holds
    -- angle in = angle out: v'.n = -v.n and v'.t = v.t for the segment hit
    (Segment.cpp:40), speed unchanged."""
    c = cases.flow_case(geometry, mach=5.0, kn=0.01, nx=80, ny=40, ppc=10.0, half_angle_deg=25.0, n_points=128)
    p = c["params"]
    p.wall_bc = abi.BC_SPECULAR
    p.tau = p.tau * 8
    body = c["body"]
    oc, info, tables = O.case_from_params(p, body)
    rng = np.random.default_rng(11 if geometry == "cylinder" else 12)
    seg6 = tables["seg"].reshape(-1, 6)
    # start within reach of the body: a band around its bounding box
    n = 40000
    lo = np.maximum(body.min(0) - 0.05, 0.0)
    hi = np.minimum(body.max(0) + 0.05, [p.L1, p.L2])
    xs, ys = [], []
    while sum(len(a) for a in xs) < n:
        px = lo[0] + rng.random(4 * n) * (hi[0] - lo[0])
        py = lo[1] + rng.random(4 * n) * (hi[1] - lo[1])
        ok = ~_inside(px, py, body) & (_seg_dist(px, py, seg6.astype(np.float64)).min(1) > 1e-6)
        xs.append(px[ok]); ys.append(py[ok])
    x = np.concatenate(xs)[:n].astype(np.float32)
    y = np.concatenate(ys)[:n].astype(np.float32)
    th = rng.random(n) * 2 * np.pi
    spd = (0.2 + 5.8 * rng.random(n)) * info.margin / info.dt
    parts = dict(x=x, y=y, vx=(spd * np.cos(th)).astype(np.float32), vy=(spd * np.sin(th)).astype(np.float32),
                 vz=rng.normal(0, 300, n).astype(np.float32), pid=np.arange(n, dtype=np.uint32))
    s, ev, keep, seg = _compare_with_brute_force(oc, info, tables, p, parts, geometry)
    single = np.array([len(e) == 1 and e[0][0] >= 4 for e in ev]) & keep
    assert single.sum() > 1500
    k = np.array([e[0][1] if len(e) == 1 else 0 for e in ev])[single]
    nx_, ny_ = seg[k, 4], seg[k, 5]
    v0 = np.stack([parts["vx"][single], parts["vy"][single]], 1).astype(np.float64)
    v1 = np.stack([s["vx"][single], s["vy"][single]], 1).astype(np.float64)
    sp = np.hypot(v0[:, 0], v0[:, 1])
    vn0 = v0[:, 0] * nx_ + v0[:, 1] * ny_
    vn1 = v1[:, 0] * nx_ + v1[:, 1] * ny_
    vt0 = -v0[:, 0] * ny_ + v0[:, 1] * nx_
    vt1 = -v1[:, 0] * ny_ + v1[:, 1] * nx_
    assert np.all(vn0 < 0) and np.all(vn1 > 0)
    np.testing.assert_allclose(vn1, -vn0, rtol=0, atol=2e-6 * sp.max())
    np.testing.assert_allclose(vt1, vt0, rtol=0, atol=2e-6 * sp.max())
    np.testing.assert_allclose(np.hypot(v1[:, 0], v1[:, 1]), sp, rtol=2e-6)


def test_diffuse_body_segment_emission_moments():
    """Particles aimed along -n at the wedge's upper face (a segment at
    20 deg to the axes) hit it half-way through the step and are re-emitted
    from the wall Maxwellian at T_w: in the segment's frame <v_n> =
    sqrt(pi k T_w / 2m), <v_n^2> = 2 k T_w / m, tangential and z components
    N(0, k T_w / m), all three independent; the particle then flies the
    remaining half step from the hit point."""
    c = cases.flow_case("wedge", mach=5.0, kn=0.01, tw_ratio=2.5, nx=80, ny=40, ppc=10.0, half_angle_deg=20.0)
    p, body = c["params"], c["body"]
    Tw = p.T_wall
    oc, info, tables = O.case_from_params(p, body)
    seg = tables["seg"].reshape(-1, 6).astype(np.float64)
    ax, ay, bx, by, nx_, ny_ = seg[2]                      # upper face: (apex + L, cy + h) -> apex
    assert ny_ > 0 and nx_ < 0
    n = 200000
    rng = np.random.default_rng(3)
    dt = np.float64(np.float32(info.dt))
    u = 0.2 + 0.6 * rng.random(n)                           # point on the face
    hx, hy = ax + u * (bx - ax), ay + u * (by - ay)
    v_in = 400.0 + 400.0 * rng.random(n)                     # speed along -n
    x = (hx + nx_ * v_in * dt / 2).astype(np.float32)
    y = (hy + ny_ * v_in * dt / 2).astype(np.float32)
    parts = dict(x=x, y=y, vx=(-nx_ * v_in).astype(np.float32), vy=(-ny_ * v_in).astype(np.float32),
                 vz=np.zeros(n, np.float32), pid=np.arange(n, dtype=np.uint32))
    s = {k: v.copy() for k, v in parts.items()}
    dead, err = oc.move(s, 9)
    assert err == 0 and not dead.any()
    vx, vy, vz = (s[k].astype(np.float64) for k in ("vx", "vy", "vz"))
    vn = vx * nx_ + vy * ny_
    vt = -vx * ny_ + vy * nx_
    sig2 = K * Tw / M
    assert np.all(vn > 0)
    mean_vn = math.sqrt(math.pi * K * Tw / (2 * M))
    sd_vn = math.sqrt(2 * sig2 - mean_vn ** 2)
    assert _z(vn.mean() - mean_vn, sd_vn / math.sqrt(n)) < 4
    assert _z(np.mean(vn * vn) - 2 * sig2, 2 * sig2 / math.sqrt(n)) < 4
    for v in (vt, vz):
        assert _z(v.mean(), math.sqrt(sig2 / n)) < 4
        assert _z(np.mean(v * v) - sig2, sig2 * math.sqrt(2 / n)) < 4
    for a, b in ((vn, vt), (vn, vz), (vt, vz), (vn, np.abs(vt))):
        assert abs(np.corrcoef(a, b)[0, 1]) < 4 / math.sqrt(n)
    # the remaining half step from the hit point with the new velocity
    t_hit = (np.float32(v_in) * 0 + 1) * dt / 2
    ex_x = hx + vx * t_hit
    ex_y = hy + vy * t_hit
    np.testing.assert_allclose(s["x"], ex_x, rtol=0, atol=1e-5 * p.L1)
    np.testing.assert_allclose(s["y"], ex_y, rtol=0, atol=1e-5 * p.L1)


# ---------------------------------------------------------------- vertices
@pytest.mark.parametrize("half", [10.0, 17.0])
def test_no_trap_at_an_acute_apex(half):
    """Particles just upstream of a wedge's apex, between the lines of its
    two faces (within the faces' end overlap, a few um), with diffuse walls:
    none needs the 64-event limit, and none ends inside the body.  This is
    the region where a particle was re-emitted from one face's overlap into
    the other's over and over (seen in batched sweep runs)."""
    c = cases.flow_case("wedge", mach=5.0, kn=0.02, tw_ratio=3.0, nx=128, ny=64, ppc=30.0, half_angle_deg=half)
    p, body = c["params"], c["body"]
    oc, info, tables = O.case_from_params(p, body)
    ax, ay = body[0]
    a = math.radians(half)
    n = 20000
    rng = np.random.default_rng(1)
    r = rng.uniform(1e-8, 1e-5 * np.hypot(*(body[1] - body[0])), n)
    th = math.pi + rng.uniform(-a, a, n)
    spd = rng.uniform(100, 3000, n)
    ph = rng.uniform(0, 2 * math.pi, n)
    s = dict(x=(ax + r * np.cos(th)).astype(np.float32), y=(ay + r * np.sin(th)).astype(np.float32),
             vx=(spd * np.cos(ph)).astype(np.float32), vy=(spd * np.sin(ph)).astype(np.float32),
             vz=np.zeros(n, np.float32), pid=np.arange(n, dtype=np.uint32))
    dead, err = oc.move(s, 7)
    assert err == 0
    assert not _inside(s["x"].astype(np.float64), s["y"].astype(np.float64), body).any()


@pytest.mark.parametrize("geometry", ["cylinder", "wedge"])
def test_rays_through_vertices_do_not_enter_the_body(geometry):
    """Flights aimed within 1e-7 m of a polygon vertex, from outside, in
    directions spread over the body's interior angle there: every one is
    reflected, none ends inside the body (the segment ends overlap so that a
    ray through a shared vertex cannot pass between two rounded tests)."""
    c = cases.flow_case(geometry, mach=5.0, kn=0.01, nx=80, ny=40, ppc=10.0, half_angle_deg=15.0, n_points=64)
    p, body = c["params"], c["body"]
    oc, info, tables = O.case_from_params(p, body)
    rng = np.random.default_rng(5)
    nv = body.shape[0]
    per = 2000
    xs, ys, vxs, vys = [], [], [], []
    dt = float(np.float32(info.dt))
    for k in range(nv):
        V = body[k]
        a_in = body[k - 1] - V                  # edges leaving V along the outline
        b_in = body[(k + 1) % nv] - V
        ta = math.atan2(a_in[1], a_in[0])
        tb = math.atan2(b_in[1], b_in[0])
        span = (ta - tb) % (2 * math.pi)        # interior angle, counter-clockwise from b to a
        d = tb + span * rng.uniform(0.02, 0.98, per)   # directions into the body at V
        spd = rng.uniform(500, 2000, per)
        aim = V + rng.normal(0, 1e-7, (per, 2))
        t_hit = dt * rng.uniform(0.2, 0.8, per)
        xs.append(aim[:, 0] - np.cos(d) * spd * t_hit)
        ys.append(aim[:, 1] - np.sin(d) * spd * t_hit)
        vxs.append(np.cos(d) * spd)
        vys.append(np.sin(d) * spd)
    x, y = np.concatenate(xs), np.concatenate(ys)
    ok = ~_inside(x, y, body)
    n = int(ok.sum())
    s = dict(x=x[ok].astype(np.float32), y=y[ok].astype(np.float32), vx=np.concatenate(vxs)[ok].astype(np.float32),
             vy=np.concatenate(vys)[ok].astype(np.float32), vz=np.zeros(n, np.float32),
             pid=np.arange(n, dtype=np.uint32))
    v0 = (s["vx"].copy(), s["vy"].copy())
    dead, err = oc.move(s, 3)
    assert err == 0 and n > 0.9 * per * nv
    assert not _inside(s["x"].astype(np.float64), s["y"].astype(np.float64), body).any()
    assert np.mean((s["vx"] != v0[0]) | (s["vy"] != v0[1])) > 0.999
