"""Seeded synthetic particle sets for the direct-summation workloads.

All models are in Henon units (G = M = 1, E = -1/4) unless stated otherwise and are
returned as float64 numpy arrays in the centre-of-mass frame.

* ``plummer``          Aarseth, Henon & Wielen (1974) sampling of the Plummer sphere.
* ``king``             King (1966) model, W0 given, from the isotropic lowered-isothermal
                       distribution function (ODE solve of Poisson's equation).
* ``plummer_bh``       Plummer stars plus a central point mass at rest.
* ``ensemble_sizes``   cluster sizes log-uniform in [lo, hi].
* ``neighbour_h2``     per-particle neighbour radius^2 giving ~k expected neighbours.
* ``drand48_block``    the reference test's input recipe (testCodes/src/test_gravity_block.cpp:40-60)
                       regenerated with the POSIX drand48 generator.
"""
from __future__ import annotations

import math

import numpy as np

PLUMMER_SCALE = 3.0 * math.pi / 16.0     # Plummer a in Henon units


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _iso(rng, n):
    """Unit vectors uniform on the sphere."""
    z = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)


def _com(m, x, v):
    mt = m.sum()
    x = x - (m[:, None] * x).sum(0) / mt
    v = v - (m[:, None] * v).sum(0) / mt
    return x, v


def plummer(n, seed=1, mtot=1.0, rmax_frac=0.999):
    """Plummer sphere, n equal masses summing to mtot, Henon units for mtot = 1.
    Radii are drawn from the enclosed-mass fraction X in (0, rmax_frac)."""
    rng = _rng(seed)
    m = np.full(n, mtot / n)
    X = rng.uniform(0.0, rmax_frac, n)
    X = np.maximum(X, 1e-300)
    r = 1.0 / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    x = r[:, None] * _iso(rng, n)
    # speed fraction q from g(q) = q^2 (1 - q^2)^3.5 by rejection (AHW74)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        a = rng.uniform(0.0, 1.0, todo.size)
        b = rng.uniform(0.0, 0.1, todo.size)
        ok = b < a * a * (1.0 - a * a) ** 3.5
        q[todo[ok]] = a[ok]
        todo = todo[~ok]
    vesc = math.sqrt(2.0) * (1.0 + r * r) ** -0.25
    v = (q * vesc)[:, None] * _iso(rng, n)
    # Plummer units -> Henon units (scale a = 3 pi / 16)
    x *= PLUMMER_SCALE
    v /= math.sqrt(PLUMMER_SCALE)
    if mtot != 1.0:
        v *= math.sqrt(mtot)
    x, v = _com(m, x, v)
    return m, x, v


def plummer_density(r, mtot=1.0):
    a = PLUMMER_SCALE
    return 3.0 * mtot / (4.0 * math.pi * a ** 3) * (1.0 + (r / a) ** 2) ** -2.5


def plummer_potential(r, mtot=1.0):
    return -mtot / np.sqrt(r * r + PLUMMER_SCALE ** 2)


def plummer_bh(n, seed=5, mbh=0.01):
    """n Plummer stars of total mass 1 - mbh plus a point mass mbh at rest at the
    origin (returned last)."""
    m, x, v = plummer(n, seed=seed, mtot=1.0 - mbh)
    m = np.concatenate([m, [mbh]])
    x = np.concatenate([x, np.zeros((1, 3))])
    v = np.concatenate([v, np.zeros((1, 3))])
    return m, x, v


def _king_rho(W):
    """Dimensionless King density for W >= 0."""
    from scipy.special import erf
    W = np.maximum(W, 0.0)
    return np.exp(W) * erf(np.sqrt(W)) - np.sqrt(4.0 * W / math.pi) * (1.0 + 2.0 * W / 3.0)


def king_profile(W0):
    """Solve W'' + 2 W'/r = -9 rho(W)/rho(W0) (King radius units).  Returns
    (r, W, M) sampled on a fine grid out to the tidal radius."""
    from scipy.integrate import solve_ivp
    rho0 = float(_king_rho(np.array([W0]))[0])

    def rhs(r, y):
        W, dW, M = y
        rho = float(_king_rho(np.array([W]))[0]) / rho0
        d2 = -9.0 * rho - (2.0 / r) * dW
        return [dW, d2, 4.0 * math.pi * r * r * rho]

    def tidal(r, y):
        return y[0]
    tidal.terminal = True
    tidal.direction = -1
    r0 = 1e-6
    y0 = [W0 - 1.5 * r0 * r0, -3.0 * r0, 4.0 * math.pi / 3.0 * r0 ** 3]
    sol = solve_ivp(rhs, (r0, 1e4), y0, events=tidal, max_step=0.01, rtol=1e-10, atol=1e-12,
                    dense_output=True)
    rt = sol.t_events[0][0] if sol.t_events[0].size else sol.t[-1]
    rg = np.concatenate([[0.0], np.geomspace(r0, rt, 20000)])
    Y = sol.sol(np.clip(rg, r0, rt))
    W = np.where(rg == 0.0, W0, np.maximum(Y[0], 0.0))
    M = np.where(rg == 0.0, 0.0, Y[2])
    return rg, W, M


def king(n, W0=6.0, seed=3):
    """King model with central potential W0, n equal masses, Henon units."""
    rng = _rng(seed)
    rg, Wg, Mg = king_profile(W0)
    Mt = Mg[-1]
    u = rng.uniform(0.0, 1.0, n) * Mt
    r = np.interp(u, Mg, rg)
    W = np.interp(r, rg, Wg)
    # speed in units of sigma: p(v) ~ v^2 (exp(W - v^2/2) - 1) on [0, sqrt(2W)]
    v = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        Wt = W[todo]
        vmax = np.sqrt(2.0 * Wt)
        vv = rng.uniform(0.0, 1.0, todo.size) * vmax
        f = vv * vv * (np.exp(Wt - 0.5 * vv * vv) - 1.0)
        fmax = vmax * vmax * np.exp(Wt)  # loose bound
        ok = rng.uniform(0.0, 1.0, todo.size) * fmax < f
        v[todo[ok]] = vv[ok]
        todo = todo[~ok]
    x = r[:, None] * _iso(rng, n)
    vel = v[:, None] * _iso(rng, n)
    # Model units: G = sigma = r_King = 1, so the central density is rho0 = 9 / (4 pi) and
    # the true enclosed mass is rho0 * Mg.  Potential energy from the profile,
    # U = -int G M(r) dM / r, in units of rho0^2.
    rho0 = 9.0 / (4.0 * math.pi)
    dM = np.diff(Mg)
    rm = 0.5 * (rg[1:] + rg[:-1])
    Mm = 0.5 * (Mg[1:] + Mg[:-1])
    Ug = -np.sum(Mm * dM / np.maximum(rm, 1e-300))
    # Henon units: mass scale mu = 1 / (rho0 Mt); length scale s_L with U' = -1/2,
    # s_L = -2 U mu^2 = -2 Ug / Mt^2; velocity scale sqrt(mu / s_L).
    m = np.full(n, 1.0 / n)
    s_l = -2.0 * Ug / (Mt * Mt)
    x = x * s_l
    vel = vel / math.sqrt(rho0 * Mt * s_l)
    x, vel = _com(m, x, vel)
    return m, x, vel


def ensemble_sizes(nclusters=512, lo=256, hi=16384, seed=4):
    """Cluster sizes floor(exp(U(ln lo, ln hi)))."""
    rng = _rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), nclusters))).astype(np.int64)


def neighbour_h2(x, mean_ngb=32.0, mtot=1.0):
    """Per-particle h^2 with (4 pi / 3) h^3 n(r) = mean_ngb, n(r) = N rho_Plummer(r) / mtot."""
    n = x.shape[0]
    r = np.sqrt((x * x).sum(1))
    dens = n * plummer_density(r, mtot) / mtot
    h3 = 3.0 * mean_ngb / (4.0 * math.pi * dens)
    return h3 ** (2.0 / 3.0)


class Drand48:
    """POSIX drand48: X_{k+1} = (a X_k + c) mod 2^48, output X / 2^48, default seed
    state 0x1234ABCD330E (glibc).  Reproduces the reference test inputs bit-exactly."""

    A = 0x5DEECE66D
    C = 0xB
    M = (1 << 48) - 1

    def __init__(self, state=0x1234ABCD330E):
        self.x = state

    def __call__(self):
        self.x = (self.A * self.x + self.C) & self.M
        return self.x / float(1 << 48)

    def array(self, k):
        return np.array([self() for _ in range(k)])


def drand48_block(n):
    """Inputs of testCodes/src/test_gravity_block.cpp:40-60: positions and velocities in
    [0,1)^3, masses drand48()/n renormalised to 1, ids i+1, h2 = 0.1."""
    g = Drand48()
    pos = np.empty((n, 3)); vel = np.empty((n, 3)); mass = np.empty(n)
    tm = 0.0
    for i in range(n):
        pos[i, 0] = g(); pos[i, 1] = g(); pos[i, 2] = g()
        vel[i, 0] = g(); vel[i, 1] = g(); vel[i, 2] = g()
        mass[i] = 1.0 / n * g()
        tm += mass[i]
    mass *= 1.0 / tm
    ids = np.arange(1, n + 1, dtype=np.int32)
    h2 = np.full(n, 0.1)
    return mass, pos, vel, ids, h2


def block_levels(n, nlevels=20, seed=2):
    """Power-of-two timestep levels (0 = largest step), geometric occupancy, seeded."""
    rng = _rng(seed)
    p = 0.5 ** np.arange(nlevels)
    p /= p.sum()
    return rng.choice(nlevels, size=n, p=p)


def active_subsets(n, counts, seed=7):
    """Seeded uniform subsets (sorted addresses) of the given sizes."""
    rng = _rng(seed)
    return [np.sort(rng.choice(n, size=int(k), replace=False)).astype(np.int32) for k in counts]
