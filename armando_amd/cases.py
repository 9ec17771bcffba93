"""Initial conditions and parameter sets for the WCSPH step.

Each case is a dict with the particle arrays (real particles first, wall
particles after them, as in CERNsource/time_integration.f:104) and every
scalar of `arm_params`.

  dam_break(...)   2D dam break: a water column in a tank with three layers
                   of wall particles (ArmandoList/armando2Dcpu.cu:1790-1908
                   is the reference's own dam break; sizes, h, c0, jitter and
                   the hydrostatic start follow the 720k workload below).
  DAM_BREAK_720K   the 720,000-fluid-particle workload: 0.6 x 1.2 m column,
                   3.2 x 1.6 m tank, dx = 1e-3 m, jitter +-0.01 dx.
"""
import math

import numpy as np

# flags (include/armando.h)
ART_VISC, ART_HEAT, AV_VEL, DBC, GRAVITY, VISC, HEAT_SOURCE = 1, 2, 4, 8, 16, 32, 64
CONDUCTION, PERIODIC_X, GHOSTS, CONTACT = 128, 256, 512, 1024
EOS_TAIT, EOS_LINEAR = 0, 1
FP64, FP32 = 0, 1

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count):
    """`count` successive outputs of splitmix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, count):
    """Doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal01(seed, count):
    """N(0, 1) doubles by Box-Muller on pairs of splitmix64 uniforms:
    z0 = sqrt(-2 ln(1 - u0)) cos(2 pi u1), z1 = ... sin(2 pi u1), in order."""
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m).reshape(m, 2)
    r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
    t = 2.0 * math.pi * u[:, 1]
    return np.stack([r * np.cos(t), r * np.sin(t)], 1).ravel()[:count]


def grid_for(x, y, cell, margin_cells=4, top_extra=0.0):
    """Hash grid covering the particles with a margin (cell = 2h)."""
    ox = float(np.min(x)) - margin_cells * cell
    oy = float(np.min(y)) - margin_cells * cell
    nx = int(math.ceil((float(np.max(x)) - ox) / cell)) + margin_cells + 1
    ny = int(math.ceil((float(np.max(y)) + top_extra - oy) / cell)) + margin_cells + 1
    return ox, oy, nx, ny


def dam_break(col_w=0.6, col_h=1.2, tank_w=3.2, tank_h=1.6, dx=1e-3, jitter=0.01, seed=12345,
              h_fac=1.3, layers=3, precision=FP64, alpha=0.01, beta=0.0, eps=0.5,
              g=9.81, cfl=0.2, heat=True):
    """2D dam break with dynamic (density-integrating) wall particles.

    Fluid: x = (i+1/2)dx + U(-j,j)dx, y = (k+1/2)dx + U(-j,j)dx for i < col_w/dx,
    k < col_h/dx; the uniform draws come from splitmix64(seed), consumed row by
    row (k outer, i inner), x before y.  Walls: `layers` rows of particles at
    spacing dx outside [0, tank_w] x [0, tank_h] (bottom, left and right; the
    bottom rows run under the side walls).  Hydrostatic start
    rho = rho0 (1 + rho0 g (col_h - y)/B)^(1/7), B = rho0 c0^2/7,
    c0 = 10 sqrt(g col_h); dt = cfl h / c0.
    """
    ni = int(round(col_w / dx))
    nk = int(round(col_h / dx))
    nf = ni * nk
    I, K = np.meshgrid(np.arange(ni), np.arange(nk))
    I = I.ravel()
    K = K.ravel()
    x = (I + 0.5) * dx
    y = (K + 0.5) * dx
    if jitter:
        u = uniform01(seed, 2 * nf).reshape(nf, 2)
        x = x + (2.0 * u[:, 0] - 1.0) * jitter * dx
        y = y + (2.0 * u[:, 1] - 1.0) * jitter * dx
    nxw = int(round(tank_w / dx))
    nyw = int(round(tank_h / dx))
    wx, wy = [], []
    for L in range(layers):  # bottom, spanning the side walls
        xs = (np.arange(-layers, nxw + layers) + 0.5) * dx
        wx.append(xs)
        wy.append(np.full(xs.size, -(L + 0.5) * dx))
    for L in range(layers):  # left and right
        ys = (np.arange(0, nyw) + 0.5) * dx
        wx.append(np.full(ys.size, -(L + 0.5) * dx))
        wy.append(ys)
        wx.append(np.full(ys.size, tank_w + (L + 0.5) * dx))
        wy.append(ys)
    wx = np.concatenate(wx)
    wy = np.concatenate(wy)
    rho0 = 1000.0
    c0 = 10.0 * math.sqrt(g * col_h)
    B = rho0 * c0 * c0 / 7.0
    h = h_fac * dx

    def hydro(yy):
        return rho0 * (1.0 + rho0 * g * (col_h - yy) / B) ** (1.0 / 7.0)

    rho_f = hydro(y)
    under = (wx <= col_w) & (wy <= col_h)
    rho_w = np.where(under, hydro(np.minimum(wy, col_h)), rho0)
    X = np.concatenate([x, wx])
    Y = np.concatenate([y, wy])
    n = X.size
    cell = 2.0 * h
    ox, oy, nxg, nyg = grid_for(X, Y, cell, top_extra=0.25 * tank_h)
    flags = DBC | GRAVITY | AV_VEL | (ART_VISC if alpha or beta else 0) | (ART_HEAT if heat else 0)
    return dict(
        name="dam_break", n=n, np=nf, nv=n - nf,
        x=X, y=Y, vx=np.zeros(n), vy=np.zeros(n), rho=np.concatenate([rho_f, rho_w]),
        u=np.zeros(n),
        precision=precision, flags=flags, pa_sph=0, density_method=0,
        kernel=0, h=h, mass=rho0 * dx * dx, rho0=rho0, c0=c0, pmin=0.0,
        av_alpha=alpha, av_beta=beta, av_eta=0.1,
        ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=eps, gx=0.0, gy=-g, dt=cfl * h / c0,
        ox=ox, oy=oy, cell=cell, nx=nxg, ny=nyg,
    )


def c_path_dam_state(state, np_real, precision=FP64, alpha=0.01, eps=0.5, cfl=0.2):
    """A case from a state of the reference's C dam break (ArmandoList
    initDamBreak, armando2Dcpu.cu:1790-1908), as armando_amd.io.read_c and
    io.walls_last give it (fluid first): its material constants (rho0 = 1000,
    c0 = 50, the mass and h of the state's info columns) and its fixed grid
    (origin (-1, -1), cells of 2h over [-1, 6] x [-1, 4], nX = (int)(7 / 2h)
    + 1), stepped by this boundary's step with dynamic walls, gravity,
    Monaghan viscosity, art_heat and XSPH.  Density starts at rho0 and the
    energy at 0, as initDamBreak sets them (the state files carry no field)."""
    rho0, c0 = 1000.0, 50.0
    h = float(state["h"][0])
    mass = float(state["mass"][0])
    cell = 2.0 * h
    n = len(state["x"])
    flags = DBC | GRAVITY | AV_VEL | ART_VISC | ART_HEAT
    return dict(
        name="c_path_dam_state", n=n, np=int(np_real), nv=n - int(np_real),
        x=np.ascontiguousarray(state["x"], np.float64),
        y=np.ascontiguousarray(state["y"], np.float64),
        vx=np.ascontiguousarray(state["vx"], np.float64),
        vy=np.ascontiguousarray(state["vy"], np.float64),
        rho=np.full(n, rho0), u=np.zeros(n),
        precision=precision, flags=flags, pa_sph=0, density_method=0, kernel=0, h=h,
        mass=mass, rho0=rho0, c0=c0, pmin=0.0, av_alpha=alpha, av_beta=0.0, av_eta=0.1,
        ah_g1=0.1, ah_g2=1.0, ah_eta=0.1, xsph_eps=eps, gx=0.0, gy=-9.81, dt=cfl * h / c0,
        ox=-1.0, oy=-1.0, cell=cell, nx=int(7.0 / cell) + 1, ny=int(5.0 / cell) + 1,
    )


def dam_break_720k(precision=FP64):
    """The 720,000-fluid-particle dam break (seed 12345)."""
    return dam_break(precision=precision)


def dam_break_small(scale=0.05, precision=FP64, seed=12345, jitter=0.01, dx=1e-3):
    """Same physics at a reduced size: column/tank dimensions times `scale`."""
    return dam_break(col_w=0.6 * scale, col_h=1.2 * scale, tank_w=3.2 * scale,
                     tank_h=1.6 * scale, dx=dx, precision=precision, seed=seed, jitter=jitter)


def ghost_tank(col_w=0.03, col_h=0.06, tank_w=0.16, dx=1e-3, jitter=0.01, seed=12345,
               h_fac=1.3, spacing=0.5, precision=FP64, alpha=0.01, g=9.81, cfl=0.2):
    """Dam break against virtual walls with mirror ghosts (ghost.f).

    The fluid column is the dam break's (hydrostatic start, same jitter).
    The tank is one row of virtual particles at `spacing` dx along the
    bottom (y = 0) and the sides (x = 0, x = tank_w, up to 2 col_h), each
    carrying its inward unit normal in v, as the reference's virtual ring
    does (CERNsource/test_cases.f:185-189); rho = rho0, p = 0, u = 0.  No
    dynamic boundary: the walls act through the ghosts and their own pairs.
    """
    ni = int(round(col_w / dx))
    nk = int(round(col_h / dx))
    nf = ni * nk
    I, K = np.meshgrid(np.arange(ni), np.arange(nk))
    x = (I.ravel() + 0.5) * dx
    y = (K.ravel() + 0.5) * dx
    if jitter:
        u = uniform01(seed, 2 * nf).reshape(nf, 2)
        x = x + (2.0 * u[:, 0] - 1.0) * jitter * dx
        y = y + (2.0 * u[:, 1] - 1.0) * jitter * dx
    ds = spacing * dx
    nb_ = int(round(tank_w / ds))
    ns_ = int(round(2.0 * col_h / ds))
    bx = np.arange(nb_ + 1) * ds
    sy = np.arange(1, ns_ + 1) * ds
    wx = np.concatenate([bx, np.zeros(ns_), np.full(ns_, tank_w)])
    wy = np.concatenate([np.zeros(nb_ + 1), sy, sy])
    nxn = np.concatenate([np.zeros(nb_ + 1), np.ones(ns_), -np.ones(ns_)])
    nyn = np.concatenate([np.ones(nb_ + 1), np.zeros(ns_), np.zeros(ns_)])
    rho0 = 1000.0
    c0 = 10.0 * math.sqrt(g * col_h)
    B = rho0 * c0 * c0 / 7.0
    h = h_fac * dx
    rho_f = rho0 * (1.0 + rho0 * g * (col_h - y) / B) ** (1.0 / 7.0)
    X = np.concatenate([x, wx])
    Y = np.concatenate([y, wy])
    n = X.size
    cell = 2.0 * h
    ox, oy, nxg, nyg = grid_for(X, Y, cell)
    return dict(
        name="ghost_tank", n=n, np=nf, nv=n - nf,
        x=X, y=Y, vx=np.concatenate([np.zeros(nf), nxn]),
        vy=np.concatenate([np.zeros(nf), nyn]),
        rho=np.concatenate([rho_f, np.full(n - nf, rho0)]), u=np.zeros(n),
        precision=precision, flags=GHOSTS | GRAVITY | AV_VEL | ART_VISC | ART_HEAT,
        pa_sph=0, density_method=0, kernel=0, h=h, mass=rho0 * dx * dx, rho0=rho0, c0=c0,
        pmin=-1e30, av_alpha=alpha, av_beta=0.0, av_eta=0.1, ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=0.5, gx=0.0, gy=-g, dt=cfl * h / c0,
        ox=ox, oy=oy, cell=cell, nx=nxg, ny=nyg,
    )


def contact_tank(col_w=0.02, col_h=0.03, tank_w=0.05, dx=1e-3, gap=1.0, precision=FP64, **kw):
    """The ghost tank's column and walls with the contact displacement of
    FORTRANsource/external_force.f90 as the wall model: the fluid lattice
    starts `gap` dx from the bottom and left walls, inside the contact
    distance h / hdr (hdr = 1.2 of FORTRANsource/options.inc), so the walls
    push from the first step; the virtual particles carry their inward unit
    normals in v and take part in the pair sums as in the ghost tank."""
    c = ghost_tank(col_w=col_w, col_h=col_h, tank_w=tank_w, dx=dx, precision=precision, **kw)
    nf = c["np"]
    c["x"] = c["x"].copy()
    c["y"] = c["y"].copy()
    c["x"][:nf] += (gap - 0.5) * dx
    c["y"][:nf] += (gap - 0.5) * dx
    c["flags"] = (c["flags"] & ~GHOSTS) | CONTACT
    c["contact_hdr"] = 1.2
    c["name"] = "contact_tank"
    return c


def box_lattice(nx=20, ny=15, dx=0.01, jitter=0.05, seed=7, vel=0.0, precision=FP64, h_fac=1.3):
    """Periodic-free lattice block of real particles only (no walls), with
    jittered positions and seeded velocities, for routine-level parity."""
    I, K = np.meshgrid(np.arange(nx), np.arange(ny))
    I = I.ravel()
    K = K.ravel()
    n = I.size
    r = uniform01(seed, 6 * n).reshape(n, 6)
    x = (I + 0.5) * dx + (2 * r[:, 0] - 1) * jitter * dx
    y = (K + 0.5) * dx + (2 * r[:, 1] - 1) * jitter * dx
    vx = vel * (2 * r[:, 2] - 1)
    vy = vel * (2 * r[:, 3] - 1)
    rho = 1000.0 * (1.0 + 0.01 * (2 * r[:, 4] - 1))
    u = 1.0 + 0.5 * r[:, 5]
    h = h_fac * dx
    cell = 2.0 * h
    ox, oy, nxg, nyg = grid_for(x, y, cell)
    return dict(
        name="box", n=n, np=n, nv=0, x=x, y=y, vx=vx, vy=vy, rho=rho, u=u,
        precision=precision, flags=ART_VISC | ART_HEAT | AV_VEL, pa_sph=0, density_method=0,
        kernel=0, h=h, mass=1000.0 * dx * dx, rho0=1000.0, c0=10.0, pmin=-1e30,
        av_alpha=1.0, av_beta=1.0, av_eta=0.1, ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=0.3, gx=0.0, gy=0.0, dt=1e-5, ox=ox, oy=oy, cell=cell, nx=nxg, ny=nyg,
    )


def shear_cavity(m=41, dx=2.5e-5, lid=-1e-3, precision=FP64, h_fac=1.5, nwall=80, gap=1.0,
                 eos_p0=-0.1, layers=1):
    """Shear-driven cavity at Re ~ 1 (CERNsource/test_cases.f:577-761 with
    the options of CERNsource/sph.f:298-316), on an m x m fluid lattice of
    spacing dx whose outer rows lie `gap` dx inside the box walls; one layer
    of nwall wall particles per side (`layers` > 1 adds rings outside),
    evenly spaced along the perimeter with each corner once, the top side
    moving at `lid` in x.  rho0 = 1000, dynamic viscosity 1e-3 Pa s, linear
    EOS p = c^2 rho + eos_p0 with c = 0.01 m/s (eos_p0 = -c^2 rho0 is the
    reference's eos_poly with a1 = 0.1), summation density, pa_sph 1, XSPH
    eps 0.3 (av_vel.f:35), Wendland C2 with h = h_fac dx, dt = 5e-5 s.
    Walls are frozen (no density or energy integration) and every particle
    has the mass rho0 dx^2."""
    L = (m - 1 + 2 * gap) * dx
    I, K = np.meshgrid(np.arange(m), np.arange(m))
    x = (I.ravel() + gap) * dx
    y = (K.ravel() + gap) * dx
    nf = x.size
    wx, wy, wvx = [], [], []
    for k in range(layers):
        d = k * (L / nwall)          # layer k runs d outside the box
        nk = nwall + 8 * k
        t = np.arange(nk) * ((L + 2 * d) / nk)
        wx += [t - d, np.full(nk, L + d), L + d - t, np.full(nk, -d)]
        wy += [np.full(nk, -d), t - d, np.full(nk, L + d), L + d - t]
        wvx += [np.zeros(nk), np.zeros(nk), np.full(nk, lid), np.zeros(nk)]
    wx = np.concatenate(wx)
    wy = np.concatenate(wy)
    wvx = np.concatenate(wvx)
    X = np.concatenate([x, wx])
    Y = np.concatenate([y, wy])
    n = X.size
    vx = np.concatenate([np.zeros(nf), wvx])
    h = h_fac * dx
    cell = 2.0 * h
    # the grid leaves room for the few particles that leak through the
    # single wall layer at the lid corners
    ox, oy, nxg, nyg = grid_for(X, Y, cell, margin_cells=40)
    return dict(
        name="shear_cavity", n=n, np=nf, nv=n - nf, x=X, y=Y, vx=vx, vy=np.zeros(n),
        rho=np.full(n, 1000.0), u=np.zeros(n),
        precision=precision, flags=VISC | AV_VEL, pa_sph=1, density_method=1, kernel=0,
        h=h, mass=1000.0 * dx * dx, rho0=1000.0, c0=0.01, pmin=-1e30,
        av_alpha=0.0, av_beta=0.0, av_eta=0.1, ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=0.3, gx=0.0, gy=0.0, dt=5e-5, ox=ox, oy=oy, cell=cell, nx=nxg, ny=nyg,
        visc_eta=1e-3, eos=EOS_LINEAR, eos_p0=eos_p0,
    )


def dam_break_adjoint_100k(precision=FP64, seed=777, vstd=0.01):
    """The 100k adjoint workload: a 0.3 x 0.6 m column, dx = 1.34e-3 m, in the
    3.2 x 1.6 m tank, with initial velocities N(0, vstd) per fluid particle
    (normal01(seed): vx of every fluid particle, then vy).  Returns (case,
    objective ids): the wall particles of the left wall (x < 0) below
    y = 0.1 m."""
    c = dam_break(col_w=0.3, col_h=0.6, dx=1.34e-3, precision=precision)
    nf = c["np"]
    z = normal01(seed, 2 * nf)
    c["vx"][:nf] = vstd * z[:nf]
    c["vy"][:nf] = vstd * z[nf:]
    obj = np.nonzero((c["x"] < 0) & (c["y"] < 0.1))[0].astype(np.int32)
    return c, obj


def with_heat_source(case, q0, x, y, sigma):
    """A copy of `case` with the volumetric Gaussian heat source
    q0 exp(-|x - (x, y)|^2 / (2 sigma^2)) (W/m^3) on its real particles."""
    c = dict(case)
    c["flags"] = case["flags"] | HEAT_SOURCE
    c["hs_q0"], c["hs_x"], c["hs_y"], c["hs_sigma"] = float(q0), float(x), float(y), float(sigma)
    return c


def heated_channel(length=0.4, width=0.05, dx=1e-4, layers=3, umax=0.5, rho0=1.0e4,
                   cond_k=15.0, cp=145.0, mu=1.8e-3, T0=300.0, sigma=5e-3, seed=2024,
                   h_fac=1.3, precision=FP32, cfl=0.2):
    """2D heated channel: a fluid layer of `width` between two no-slip wall
    blocks of `layers` particle rows, periodic in x with period `length`.

    Fluid on the lattice ((i+1/2) dx, (j+1/2) dx), i < length/dx,
    j < width/dx, no jitter; walls at y = -(l+1/2) dx and width + (l+1/2) dx.
    Poiseuille flow u(y) = umax (1 - ((y - width/2)/(width/2))^2) held by the
    body force gx = 2 (mu/rho0) umax / (width/2)^2 against the physical
    viscosity mu (so the start is the steady profile); Tait EOS with
    c0 = 10 umax; Cleary-Monaghan conduction (k, cp) plus artificial heat;
    u = cp T0 initially; a Gaussian heat source of width sigma at the centre
    of the channel with amplitude Q0 = 1e8 + 9e8 U, U = uniform01(seed)[0]
    (W/m^3); dt = cfl h / c0.  The periodic grid has nx = floor(length/2h)
    cells of size length/nx >= 2h."""
    ni = int(round(length / dx))
    nj = int(round(width / dx))
    L = ni * dx
    I, J = np.meshgrid(np.arange(ni), np.arange(nj))
    x = (I.ravel() + 0.5) * dx
    y = (J.ravel() + 0.5) * dx
    H = 0.5 * nj * dx
    vx = umax * (1.0 - ((y - H) / H) ** 2)
    wx, wy = [], []
    for lyr in range(layers):
        xs = (np.arange(ni) + 0.5) * dx
        wx += [xs, xs]
        wy += [np.full(ni, -(lyr + 0.5) * dx), np.full(ni, nj * dx + (lyr + 0.5) * dx)]
    X = np.concatenate([x] + wx)
    Y = np.concatenate([y] + wy)
    n, nf = X.size, x.size
    h = h_fac * dx
    nxg = int(math.floor(L / (2.0 * h)))
    cell = L / nxg
    oy = float(np.min(Y)) - 4 * cell
    nyg = int(math.ceil((float(np.max(Y)) - oy) / cell)) + 5
    c0 = 10.0 * umax
    gx = 2.0 * (mu / rho0) * umax / (H * H)
    q0 = 1e8 + 9e8 * float(uniform01(seed, 1)[0])
    flags = (DBC | GRAVITY | ART_HEAT | VISC | HEAT_SOURCE | CONDUCTION | PERIODIC_X)
    return dict(
        name="heated_channel", n=n, np=nf, nv=n - nf,
        x=X, y=Y, vx=np.concatenate([vx, np.zeros(n - nf)]), vy=np.zeros(n),
        rho=np.full(n, rho0), u=np.full(n, cp * T0),
        precision=precision, flags=flags, pa_sph=1, density_method=0,
        kernel=0, h=h, mass=rho0 * dx * dx, rho0=rho0, c0=c0, pmin=-1e30,
        av_alpha=0.0, av_beta=0.0, av_eta=0.1, ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=0.0, gx=gx, gy=0.0, dt=cfl * h / c0,
        ox=0.0, oy=oy, cell=cell, nx=nxg, ny=nyg,
        visc_eta=mu, hs_q0=q0, hs_x=0.5 * L, hs_y=H, hs_sigma=sigma,
        cond_k=cond_k, cp=cp,
    )


def heated_channel_small(length=0.02, width=0.005, dx=1e-4, precision=FP64, **kw):
    """The heated channel at a reduced size (same dx and physics)."""
    return heated_channel(length=length, width=width, dx=dx, precision=precision, **kw)


# ------------------------------------------------------------------ 3D
def grid_for3(x, y, z, cell, margin_cells=4, top_extra=0.0):
    """3D hash grid covering the particles with a margin (cell = 2h)."""
    o, n = [], []
    for a, extra in ((x, 0.0), (y, 0.0), (z, top_extra)):
        lo = float(np.min(a)) - margin_cells * cell
        o.append(lo)
        n.append(int(math.ceil((float(np.max(a)) + extra - lo) / cell)) + margin_cells + 1)
    return o[0], o[1], o[2], n[0], n[1], n[2]


def box_lattice3(nx=10, ny=9, nz=8, dx=0.01, jitter=0.1, seed=7, vel=0.0, precision=FP64,
                 h_fac=1.3):
    """3D lattice block of real particles only (no walls), jittered, with
    seeded velocities, for routine-level parity.  Draws per particle (i fastest,
    then j, then k): x, y, z jitter, vx, vy, vz, rho, u."""
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    n = I.size
    r = uniform01(seed, 8 * n).reshape(n, 8)
    x = (I + 0.5) * dx + (2 * r[:, 0] - 1) * jitter * dx
    y = (J + 0.5) * dx + (2 * r[:, 1] - 1) * jitter * dx
    z = (K + 0.5) * dx + (2 * r[:, 2] - 1) * jitter * dx
    h = h_fac * dx
    cell = 2.0 * h
    ox, oy, oz, nxg, nyg, nzg = grid_for3(x, y, z, cell)
    return dict(
        name="box3", dim=3, n=n, np=n, nv=0, x=x, y=y, z=z,
        vx=vel * (2 * r[:, 3] - 1), vy=vel * (2 * r[:, 4] - 1), vz=vel * (2 * r[:, 5] - 1),
        rho=1000.0 * (1.0 + 0.01 * (2 * r[:, 6] - 1)), u=1.0 + 0.5 * r[:, 7],
        precision=precision, flags=ART_VISC | ART_HEAT | AV_VEL, pa_sph=0, density_method=0,
        kernel=3, h=h, mass=1000.0 * dx ** 3, rho0=1000.0, c0=10.0, pmin=-1e30,
        av_alpha=1.0, av_beta=1.0, av_eta=0.1, ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=0.3, gx=0.0, gy=0.0, gz=0.0, dt=1e-5, ox=ox, oy=oy, oz=oz, cell=cell,
        nx=nxg, ny=nyg, nz=nzg,
    )


def dam_break3d(col=(0.6, 0.6, 1.2), tank=(3.2, 1.0, 1.6), dx=7.1e-3, jitter=0.01, seed=4242,
                h_fac=1.3, layers=3, precision=FP64, alpha=0.01, beta=0.0, eps=0.5, g=9.81,
                cfl=0.2, heat=True):
    """3D dam break with dynamic (density-integrating) wall particles.

    Fluid: the lattice (i+1/2, j+1/2, k+1/2) dx, i < col_x/dx, j < col_y/dx,
    k < col_z/dx, in order k outer, j, i inner, each coordinate moved by
    U(-jitter, jitter) dx (splitmix64(seed), x, y, z per particle).  Walls:
    every lattice point (i+1/2, j+1/2, k+1/2) dx with -layers <= i <
    nxw + layers, -layers <= j < nyw + layers, -layers <= k < nzw that lies
    outside the tank [0, nxw) x [0, nyw) x [0, nzw) (a box open at the top,
    `layers` thick).  Hydrostatic start in z, gravity -g along z,
    c0 = 10 sqrt(g col_z), dt = cfl h / c0, Wendland C2 3D."""
    ni, nj, nk = (int(round(c / dx)) for c in col)
    K, J, I = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    nf = I.size
    x = (I + 0.5) * dx
    y = (J + 0.5) * dx
    z = (K + 0.5) * dx
    if jitter:
        u = uniform01(seed, 3 * nf).reshape(nf, 3)
        x = x + (2.0 * u[:, 0] - 1.0) * jitter * dx
        y = y + (2.0 * u[:, 1] - 1.0) * jitter * dx
        z = z + (2.0 * u[:, 2] - 1.0) * jitter * dx
    nxw, nyw, nzw = (int(round(t / dx)) for t in tank)
    L = layers
    K, J, I = np.meshgrid(np.arange(-L, nzw), np.arange(-L, nyw + L), np.arange(-L, nxw + L),
                          indexing="ij")
    outside = (I < 0) | (I >= nxw) | (J < 0) | (J >= nyw) | (K < 0)
    wx = (I[outside] + 0.5) * dx
    wy = (J[outside] + 0.5) * dx
    wz = (K[outside] + 0.5) * dx
    rho0 = 1000.0
    col_h = col[2]
    c0 = 10.0 * math.sqrt(g * col_h)
    B = rho0 * c0 * c0 / 7.0
    h = h_fac * dx

    def hydro(zz):
        return rho0 * (1.0 + rho0 * g * (col_h - zz) / B) ** (1.0 / 7.0)

    under = (wx <= col[0]) & (wy <= col[1]) & (wz <= col_h)
    rho_w = np.where(under, hydro(np.minimum(wz, col_h)), rho0)
    X = np.concatenate([x, wx])
    Y = np.concatenate([y, wy])
    Z = np.concatenate([z, wz])
    n = X.size
    cell = 2.0 * h
    ox, oy, oz, nxg, nyg, nzg = grid_for3(X, Y, Z, cell, top_extra=0.25 * tank[2])
    flags = DBC | GRAVITY | AV_VEL | (ART_VISC if alpha or beta else 0) | (ART_HEAT if heat else 0)
    return dict(
        name="dam_break3d", dim=3, n=n, np=nf, nv=n - nf,
        x=X, y=Y, z=Z, vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n),
        rho=np.concatenate([hydro(z), rho_w]), u=np.zeros(n),
        precision=precision, flags=flags, pa_sph=0, density_method=0,
        kernel=3, h=h, mass=rho0 * dx ** 3, rho0=rho0, c0=c0, pmin=0.0,
        av_alpha=alpha, av_beta=beta, av_eta=0.1,
        ah_g1=0.1, ah_g2=1.0, ah_eta=0.1,
        xsph_eps=eps, gx=0.0, gy=0.0, gz=-g, dt=cfl * h / c0,
        ox=ox, oy=oy, oz=oz, cell=cell, nx=nxg, ny=nyg, nz=nzg,
    )


def dam_break3d_small(scale=0.1, dx=7.1e-3, precision=FP64, seed=4242, jitter=0.01):
    """The 3D dam break with column and tank dimensions times `scale`."""
    return dam_break3d(col=(0.6 * scale, 0.6 * scale, 1.2 * scale),
                       tank=(3.2 * scale, 1.0 * scale, 1.6 * scale), dx=dx,
                       precision=precision, seed=seed, jitter=jitter)


PARAM_KEYS = ("precision", "flags", "pa_sph", "density_method", "h", "mass", "rho0", "c0",
              "pmin", "av_alpha", "av_beta", "av_eta", "ah_g1", "ah_g2", "ah_eta", "xsph_eps",
              "gx", "gy", "dt", "ox", "oy", "cell", "nx", "ny", "visc_eta", "eos", "eos_p0",
              "hs_q0", "hs_x", "hs_y", "hs_sigma", "dim", "oz", "gz", "nz", "cond_k", "cp",
              "sle", "contact_hdr")
