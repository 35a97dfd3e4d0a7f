"""Case definitions: gas constants, the benchmark configurations and body
polygons.  The role of the reference's compile-time src/Param.h:5-37 and of
the geometry choice in Run::initialize (Run.cpp:45-48), as runtime values.
"""
import math

import numpy as np

from . import abi

BOLTZ = 1.380649e-23          # J/K

# argon, VHS (Bird 1994, Appendix A)
ARGON = dict(mass=6.63e-26, diam=4.17e-10, Vtl=0.81, Tref=273.0)
# nitrogen, VHS (Bird 1994, Appendix A), two rotational degrees of freedom:
# the reference's default gas (src/Param.h:7,14: the N2 mass 4.65116e-26 kg,
# zeta = 2; it keeps argon's omega 0.81 there)
NITROGEN = dict(mass=4.65e-26, diam=4.17e-10, Vtl=0.74, Tref=273.0, zeta=2)


def vhs_collision_frequency(n, T, gas=ARGON, boltz=BOLTZ):
    """Equilibrium VHS collision frequency (Bird 1994 eq. 4.64):
    nu = 4 d^2 n sqrt(pi k Tref / m) (T/Tref)^(1-omega)."""
    d, m, w, Tr = gas["diam"], gas["mass"], gas["Vtl"], gas["Tref"]
    return 4.0 * d * d * n * math.sqrt(math.pi * boltz * Tr / m) * (T / Tr) ** (1.0 - w)


def hard_sphere_density_for_mfp(lam, gas=ARGON):
    """n such that the hard-sphere mean free path 1/(sqrt(2) pi d^2 n) is lam."""
    return 1.0 / (math.sqrt(2.0) * math.pi * gas["diam"] ** 2 * lam)


def circle_polygon(n, cx, cy, r):
    """Counter-clockwise regular polygon (Circle::Initialize, Circle.cpp:10-19)."""
    a = 2.0 * math.pi * np.arange(n) / n
    return np.stack([cx + r * np.cos(a), cy + r * np.sin(a)], axis=1)


def wedge_polygon(apex_x, cy, length, half_angle_deg):
    """Counter-clockwise isosceles wedge, apex upstream at (apex_x, cy)."""
    h = length * math.tan(math.radians(half_angle_deg))
    return np.array([[apex_x, cy], [apex_x + length, cy - h], [apex_x + length, cy + h]], dtype=np.float64)


def gas_fields(gas, rot_prob=1.0):
    """make_params keywords of a gas: VHS constants and rotational degrees
    of freedom (zeta = 2 with rot_prob = 1/Z_rot for a diatomic gas)."""
    d = {k: gas[k] for k in ("mass", "diam", "Vtl", "Tref")}
    d["zeta"] = gas.get("zeta", 0)
    d["rot_prob"] = rot_prob
    return d


def sound_speed(T, gas=ARGON, boltz=BOLTZ):
    """sqrt(gamma k T / m) with gamma = (5 + zeta) / (3 + zeta)."""
    z = gas.get("zeta", 0)
    return math.sqrt((5.0 + z) / (3.0 + z) * boltz * T / gas["mass"])


def make_params(**kw):
    """dsmc_params with defaults; keyword names are the struct's fields."""
    p = abi.dsmc_params()
    p.boltz = BOLTZ
    for k, v in ARGON.items():
        setattr(p, k, v)
    p.L3 = 1.0
    p.V_y = 0.0
    p.T_wall = 300.0
    p.wall_bc = abi.BC_DIFFUSE
    for f in range(4):
        p.face_bc[f] = abi.BC_STREAM
        p.face_T[f] = 0.0
    p.Fn = 0.0
    p.seed = 20251203
    p.capacity = 0
    p.sample_start = -1
    p.zeta = 0
    p.rot_prob = 1.0
    for k, v in kw.items():
        if k in ("face_bc", "face_T"):
            arr = getattr(p, k)
            for f in range(4):
                arr[f] = v[f]
        else:
            setattr(p, k, v)
    return p


def box_equilibrium(ppc=20, n_cells=64, steps=1000, seed=20251203):
    """Plain DSMC, 2D closed box at equilibrium: argon, 1 mm x 1 mm, 64x64
    cells, n = 1e20 m^-3, T = 300 K, 20 particles per cell, specular walls,
    dt = 0.2 / nu(VHS), 1000 steps."""
    n, T = 1e20, 300.0
    nu = vhs_collision_frequency(n, T)
    p = make_params(L1=1e-3, L2=1e-3, N1=n_cells, N2=n_cells, n_inf=n, T=T, V_jet=0.0,
                    wall_bc=abi.BC_SPECULAR, face_bc=[abi.BC_SPECULAR] * 4,
                    tau=0.2 / nu, ppc=float(ppc), seed=seed, sample_start=0)
    return dict(name="box_equilibrium_64x64", params=p, body=None, steps=steps, nu=nu)


def cylinder_mach10(ppc=30, nx=400, ny=200, n_points=256, steps=20000, sample_from=10000,
                    seed=20251203, gas=ARGON, rot_prob=1.0):
    """Plain DSMC, 2D supersonic argon flow over a cut-cell cylinder: Mach 10,
    T_inf = 200 K, Kn = 0.01 on the diameter (hard-sphere mean free path),
    R = 0.1 m centred in a 0.6 m x 0.3 m domain, 400x200 cells, diffuse wall
    at 500 K, freestream inflow/outflow on all faces, 30 particles per fluid
    cell, dt = 0.3 dx / (U + 3 vmp), 20k steps, sampling from step 10k.  gas:
    another gas for the same geometry (NITROGEN: Mach 10 with gamma = 7/5)."""
    T = 200.0
    m = gas["mass"]
    a = sound_speed(T, gas)
    U = 10.0 * a
    R, D = 0.1, 0.2
    lam = 0.01 * D
    n = hard_sphere_density_for_mfp(lam, gas)
    L1, L2 = 0.6, 0.3
    dx = L1 / nx
    vmp = math.sqrt(2.0 * BOLTZ * T / m)
    dt = 0.3 * dx / (U + 3.0 * vmp)
    p = make_params(L1=L1, L2=L2, N1=nx, N2=ny, n_inf=n, T=T, V_jet=U, T_wall=500.0,
                    wall_bc=abi.BC_DIFFUSE, face_bc=[abi.BC_STREAM] * 4, tau=dt,
                    ppc=float(ppc), seed=seed, sample_start=sample_from, **gas_fields(gas, rot_prob))
    body = circle_polygon(n_points, 0.5 * L1, 0.5 * L2, R)
    return dict(name=f"cylinder_mach10_{nx}x{ny}", params=p, body=body, steps=steps, U=U,
                sample_from=sample_from)


def flow_case(geometry="cylinder", mach=10.0, kn=0.01, tw_ratio=2.5, nx=400, ny=200, ppc=30.0,
              half_angle_deg=20.0, T_inf=200.0, steps=20000, sample_from=10000, seed=20251203, n_points=256,
              gas=ARGON, rot_prob=1.0):
    """Supersonic argon flow over a cylinder (R = 0.1 m, centred) or a wedge
    (apex at x = 0.15 m, length 0.3 m or less so that its base stays 1 cm
    inside the domain) in a 0.6 m x 0.3 m domain:
    freestream inflow/outflow on all faces, diffuse wall at tw_ratio*T_inf,
    Kn on the body size (cylinder diameter, wedge length) with the
    hard-sphere mean free path, dt = 0.3 dx / (U + 3 vmp); gas: argon unless
    given (Mach on the gas's own speed of sound)."""
    m = gas["mass"]
    a = sound_speed(T_inf, gas)
    U = mach * a
    L1, L2 = 0.6, 0.3
    if geometry == "cylinder":
        size = 0.2
        body = circle_polygon(n_points, 0.5 * L1, 0.5 * L2, 0.1)
    elif geometry == "wedge":
        # 0.3 m long unless its base would come closer than 1 cm to the
        # y faces (half-angles above about 25 deg): the body stays inside
        # the domain, away from the stream faces
        size = min(0.3, (0.5 * L2 - 0.01) / math.tan(math.radians(half_angle_deg)))
        body = wedge_polygon(0.15, 0.5 * L2, size, half_angle_deg)
    else:
        raise ValueError(f"unknown geometry {geometry!r}")
    n = hard_sphere_density_for_mfp(kn * size, gas)
    dx = L1 / nx
    vmp = math.sqrt(2.0 * BOLTZ * T_inf / m)
    dt = 0.3 * dx / (U + 3.0 * vmp)
    p = make_params(L1=L1, L2=L2, N1=nx, N2=ny, n_inf=n, T=T_inf, V_jet=U, T_wall=tw_ratio * T_inf,
                    wall_bc=abi.BC_DIFFUSE, face_bc=[abi.BC_STREAM] * 4, tau=dt,
                    ppc=float(ppc), seed=seed, sample_start=sample_from, **gas_fields(gas, rot_prob))
    return dict(name=f"{geometry}_M{mach:.2f}_Kn{kn:.4f}_Tw{tw_ratio:.2f}_{nx}x{ny}", params=p, body=body,
                steps=steps, U=U, sample_from=sample_from,
                point=dict(geometry=geometry, mach=mach, kn=kn, tw_ratio=tw_ratio,
                           half_angle_deg=half_angle_deg if geometry == "wedge" else 0.0))


SWEEP_GRIDS = [(64, 32), (96, 48), (128, 64), (160, 80), (192, 96), (224, 112), (256, 128)]


def sweep(n=64, seed=20251203, particles=(2e5, 6e6), ppc=(30.0, 200.0), steps=20000, sample_from=10000,
          n_points=128):
    """The batched parameter sweep, every value drawn from `seed`: geometry
    cylinder or wedge (half-angle U[10, 30] deg), Mach U[3, 12], Kn
    log-U[0.005, 0.1], T_w/T_inf U[1, 5], particle count log-U over
    `particles` (uneven instances, about 1.7M on average), and the grid of
    64x32 ... 256x128 closest to a particles-per-cell target log-U over
    `ppc`, so that cells stay DSMC-sized whatever the particle count."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        geom = "cylinder" if rng.random() < 0.5 else "wedge"
        half = float(rng.uniform(10.0, 30.0))
        mach = float(rng.uniform(3.0, 12.0))
        kn = float(math.exp(rng.uniform(math.log(0.005), math.log(0.1))))
        tw = float(rng.uniform(1.0, 5.0))
        npart = float(math.exp(rng.uniform(math.log(particles[0]), math.log(particles[1]))))
        target = npart / float(math.exp(rng.uniform(math.log(ppc[0]), math.log(ppc[1]))))
        nx, ny = min(SWEEP_GRIDS, key=lambda g: abs(math.log(g[0] * g[1] / target)))
        wl = min(0.3, 0.14 / math.tan(math.radians(half)))
        solid = wl * wl * math.tan(math.radians(half)) if geom == "wedge" else math.pi * 0.01
        fluid_cells = nx * ny * (1.0 - solid / 0.18)
        cfg = flow_case(geom, mach, kn, tw, nx, ny, ppc=npart / fluid_cells, half_angle_deg=half,
                        steps=steps, sample_from=sample_from, seed=int(rng.integers(1, 2**62)),
                        n_points=n_points)
        cfg["index"] = i
        out.append(cfg)
    return out
