"""Shared problem set-ups (initial states, exact solutions, meshes) for tests.

The physical set-ups follow the workload descriptions: the 2-D isentropic
vortex on the periodic box [-5,5]^2 (freestream rho=1, p=1, (u,v)=(1,1),
beta=5, gamma=1.4; the same expressions as the reference's
examples/Euler/isentropic_vortex/param.in:72-96 with the freestream velocity
(1,1)), the four-quadrant 2-D Riemann problem on [0,1]^2 split at (0.8,0.8)
with transmissive (outlet) sides, and the Taylor-Green vortex on the periodic
box [0,2pi]^2 at Mach 0.1, Re 1600, Pr 0.72.
"""
from __future__ import annotations

import numpy as np

import tensum_amd as T


def vortex(x, y, t=0.0, gam=1.4, beta=5.0, box=10.0, R=1.0):
    """(T,u,v,p) of the isentropic vortex advected with (1,1) for time t."""
    xc = ((x - t) + 5.0) % box - 5.0
    yc = ((y - t) + 5.0) % box - 5.0
    r2 = xc * xc + yc * yc
    Tm = 1.0 - (gam - 1) * (beta ** 2 / (8 * gam * np.pi ** 2)) * np.exp(1 - r2)
    u = 1.0 - beta / (2 * np.pi) * yc * np.exp(0.5 * (1 - r2))
    v = 1.0 + beta / (2 * np.pi) * xc * np.exp(0.5 * (1 - r2))
    p = Tm ** (gam / (gam - 1))
    return np.stack([Tm / R, u, v, p], 1)


RIEMANN_STATES = {  # (rho, u, v, p) per quadrant, split at (0.8, 0.8)
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def riemann(x, y, split=(0.8, 0.8), pscale=(1.0, 1.0, 1.0, 1.0), R=1.0):
    """(T,u,v,p) of the four-quadrant Riemann problem; x >= split is east."""
    east = x >= split[0]
    north = y >= split[1]
    out = np.empty((x.size, 4))
    for k, (key, m) in enumerate((("NE", east & north), ("NW", ~east & north),
                                  ("SW", ~east & ~north), ("SE", east & ~north))):
        rho, u, v, p = RIEMANN_STATES[key]
        p = p * pscale[k]
        out[m] = (p / (R * rho), u, v, p)
    return out


def taylor_green(x, y, mach=0.1, gam=1.4, R=1.0):
    """(T,u,v,p) of the Taylor-Green vortex, rho0 = U0 = 1."""
    p0 = 1.0 / (gam * mach ** 2)
    u = np.sin(x) * np.cos(y)
    v = -np.cos(x) * np.sin(y)
    p = p0 + 0.25 * (np.cos(2 * x) + np.cos(2 * y))
    rho = 1.0
    return np.stack([p / (R * rho), u, v, p], 1)


def vortex_mesh(n, seed=0):
    return T.jittered_mesh(n, box=(-5.0, 5.0, -5.0, 5.0), jitter=0.2, seed=seed, periodic=True)


def riemann_mesh(n, seed=1):
    return T.jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=seed, periodic=False,
                           bc={t: ("outlet",) for t in (1, 2, 3, 4)})


def tgv_mesh(n, seed=2):
    return T.jittered_mesh(n, box=(0.0, 2 * np.pi, 0.0, 2 * np.pi), jitter=0.2, seed=seed,
                           periodic=True)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def conserved(prim, gam=1.4, R=1.0):
    """(rho, rho u, rho v, E) from (T,u,v,p)."""
    Tm, u, v, p = prim.T
    rho = p / (R * Tm)
    return np.stack([rho, rho * u, rho * v, p / (gam - 1) + 0.5 * rho * (u * u + v * v)], 1)
