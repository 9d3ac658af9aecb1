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
    """(T,u,v,p) of the isentropic vortex advected with (1,1) for time t;
    centred in the periodic box of side `box`, or not wrapped (box=None, the
    param.in expressions as written)."""
    if box is None:
        xc, yc = x - t, y - t
    else:
        xc = ((x - t) + 0.5 * box) % box - 0.5 * box
        yc = ((y - t) + 0.5 * box) % box - 0.5 * box
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


def write_part_file(mesh, directory, periodic_tags=()):
    """Writes <directory>/PARTITION/part_file_0.dat for NPART = 1 in the layout of
    readme/README.md:155-201 (gids 0-based, coordinates at 17 digits)."""
    import os
    d = os.path.join(directory, "PARTITION")
    os.makedirs(d, exist_ok=True)
    n = mesh.n_vertex
    on_b = np.zeros(n, dtype=int)
    on_b[mesh.bface_v] = 1
    lines = ["DIMENSION 2", "PARTITIONS 1", f"NODES {n} OF {n}"]
    lines += [f"{i} {mesh.x[i]:.17g} {mesh.y[i]:.17g} 0 1 0 {on_b[i]}" for i in range(n)]
    nper = mesh.periodic_ptr.size - 1 if mesh.periodic_ptr.size else 0
    lines.append(f"PERIODIC {nper}")
    for k in range(nper):
        mem = mesh.periodic_v[mesh.periodic_ptr[k]:mesh.periodic_ptr[k + 1]]
        lines.append(f"{k} {mem.size} " + " ".join(map(str, mem)) + " 1 0")
    nb = mesh.bface_tag.size
    lines.append(f"FACES {nb} OF {nb}")
    bv = mesh.bface_v.reshape(-1, 2)
    for k in range(nb):
        per = 1 if mesh.bface_tag[k] in periodic_tags else 0
        lines.append(f"{mesh.bface_tag[k]} {bv[k, 0]} {bv[k, 1]} {per}")
    cv = mesh.cell_v.reshape(-1, 3)
    lines.append(f"CELLS {cv.shape[0]} OF {cv.shape[0]}")
    lines += [f"{a} {b} {c} 0" for a, b, c in cv]
    with open(os.path.join(d, "part_file_0.dat"), "w") as f:
        f.write("\n".join(lines) + "\n")


def write_msh_strips(mesh, path, nparts):
    """Writes `mesh` as a gmsh-2.2 ASCII file for the reference's partitioner
    (grid_gen/gmsh_split.cc:15-192, grid_gen/read_gmsh.cc:16-206), split into
    `nparts` partitions by x-strips of the cell centroids: each triangle's tags
    are (1, 1, number of partitions, owner + 1, -(ghost + 1)...), its ghosts
    being every other partition that owns a cell sharing one of its vertices
    (readme/README.md:153); each boundary line's tags are (its boundary tag,
    the same as elementary tag, 1, owner of its cell + 1)."""
    n = mesh.n_vertex
    cv = mesh.cell_v.reshape(-1, 3).astype(np.int64)
    nc = cv.shape[0]
    cx = (mesh.x[cv[:, 0]] + mesh.x[cv[:, 1]] + mesh.x[cv[:, 2]]) / 3.0
    x0, x1 = float(mesh.x.min()), float(mesh.x.max())
    owner = np.minimum((nparts * (cx - x0) / (x1 - x0)).astype(np.int64), nparts - 1)
    # the smallest and largest owner among each vertex's cells
    vmin = np.full(n, nparts, dtype=np.int64)
    vmax = np.full(n, -1, dtype=np.int64)
    for j in range(3):
        np.minimum.at(vmin, cv[:, j], owner)
        np.maximum.at(vmax, cv[:, j], owner)
    # owners of a vertex's cells are contiguous strips, so a cell's ghosts are
    # the strips in [min, max] over its corners other than its own
    lo = np.minimum.reduce([vmin[cv[:, j]] for j in range(3)])
    hi = np.maximum.reduce([vmax[cv[:, j]] for j in range(3)])
    # boundary lines: the owner of the cell holding the edge
    bv = mesh.bface_v.reshape(-1, 2).astype(np.int64)
    key = lambda a, b: np.minimum(a, b) * n + np.maximum(a, b)
    bkey = key(bv[:, 0], bv[:, 1])
    bown = np.full(bv.shape[0], -1, dtype=np.int64)
    order = np.argsort(bkey)
    for j in range(3):
        k = key(cv[:, j], cv[:, (j + 1) % 3])
        pos = np.searchsorted(bkey[order], k)
        pos = np.minimum(pos, bkey.size - 1)
        hit = bkey[order][pos] == k
        bown[order[pos[hit]]] = owner[hit]
    assert (bown >= 0).all(), "boundary edge without a cell"
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(n)]
    lines += [f"{i + 1} {mesh.x[i]:.17g} {mesh.y[i]:.17g} 0" for i in range(n)]
    lines += ["$EndNodes", "$Elements", str(bv.shape[0] + nc)]
    e = 0
    for k in range(bv.shape[0]):
        e += 1
        t = int(mesh.bface_tag[k])
        lines.append(f"{e} 1 4 {t} {t} 1 {bown[k] + 1} {bv[k, 0] + 1} {bv[k, 1] + 1}")
    for c in range(nc):
        e += 1
        o = int(owner[c])
        gh = [p for p in range(int(lo[c]), int(hi[c]) + 1) if p != o]
        tags = [1, 1, 1 + len(gh), o + 1] + [-(p + 1) for p in gh]
        lines.append(f"{e} 2 {len(tags)} " + " ".join(map(str, tags)) +
                     f" {cv[c, 0] + 1} {cv[c, 1] + 1} {cv[c, 2] + 1}")
    lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return owner


VORTEX_PARAM = """grid
{{
   cell      median
}}
numeric
{{
   time_mode    unsteady
   time_scheme  ssprk3
   time_step    0.0
   cfl          0.8
   max_iter     5000
   final_time   {final_time}
   min_residue  0.0
   reconstruct
   {{
      method   minmod
   }}
   bc_scheme    weak
   liou_fix     no
   sample_list
   {{
      groups
      {{
      }}
      free_list
      {{
      }}
   }}
   rnd_file_loc    rnd_nos.dat
   rnd_per_sample  0
}}
material
{{
   gamma       1.4
   gas_const   1
   model       euler
   viscosity
   {{
      model    constant
      mu_ref   0.0
   }}
   prandtl     0.72
   flux        kepes_tecno_roe
   flux_par
   {{
      alpha      0
      beta       0
   }}
   balance
   {{
       switch_active  no
       ND             yes
       HEAT           yes
       VISC           yes
   }}
}}
constants
{{
   gam   1.4
   beta  5
}}
initial_condition
{{
   temperature  1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x,2)-pow(y,2)))
   xvelocity    1 - (beta/(2*pi))*(y)*exp(0.5*(1-pow(x,2)-pow(y,2)))
   yvelocity    1 + (beta/(2*pi))*(x)*exp(0.5*(1-pow(x,2)-pow(y,2)))
   zvelocity    0.0
   pressure     pow(1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x,2)-pow(y,2))),gam/(gam-1))
}}
boundary
{{
   1 2 3 4
   {{
      type         periodic
   }}
}}
exact_soln
{{
   available    yes
   temperature  1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x-t,2)-pow(y-t,2)))
   xvelocity    1 - (beta/(2*pi))*(y-t)*exp(0.5*(1-pow(x-t,2)-pow(y-t,2)))
   yvelocity    1 + (beta/(2*pi))*(x-t)*exp(0.5*(1-pow(x-t,2)-pow(y-t,2)))
   zvelocity    0.0
   pressure     pow(1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x-t,2)-pow(y-t,2))),gam/(gam-1))
}}
integrals
{{
}}
output
{{
   format                vtk
   frequency             100
   time_stamps           {stamps}
   use_online_stat_tool  no
   variables
   {{
       density
       vorticity
   }}
   surfaces
   {{
   }}
   restart               yes
   restart_frequency     10
   log_output            yes
   soln_output           yes
   save_global           yes
   find_error            yes
}}
"""
