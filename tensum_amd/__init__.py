"""tensum_amd -- Python host binding of libtsm, the MI355X vertex-centred
finite-volume residual + Runge-Kutta library.

This module is a thin ctypes layer over the C ABI declared in include/tsm.h.
It mirrors the reference's configuration vocabulary (flux, reconstruct and
boundary names of src/parameter.cc:114-143, :354-372 and src/bc.h:14-23) so a
caller can describe a run the way a TENSUM param.in does.  All numerics run in
the HIP library; if libtsm.so is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsm.so")

FLUX = {"kepes_tecno_roe": 0, "kepes_tecno_rusanov": 1, "kepec": 2, "roe_ec": 3,
        "kep": 4, "simple_avg": 5}
RECONSTRUCT = {"first": 0, "second": 1, "second_muscl": 2, "van_albada": 3,
               "minmod": 4, "tvb_minmod": 5, "eno2": 6}
BC = {"none": 0, "slip": 1, "noslip": 2, "farfield": 3, "inlet": 4, "outlet": 5,
      "pressure": 6, "periodic": 7}


class TsmParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("gas_const", C.c_double), ("prandtl", C.c_double),
                ("mu_ref", C.c_double), ("T_ref", C.c_double), ("T_0", C.c_double),
                ("omega", C.c_double),
                ("model", C.c_int), ("mu_model", C.c_int), ("flux", C.c_int),
                ("reconstruct", C.c_int), ("tvb_M", C.c_double),
                ("ev_alpha", C.c_double), ("ev_beta", C.c_double),
                ("nd_active", C.c_int), ("visc_active", C.c_int), ("heat_active", C.c_int),
                ("time_scheme", C.c_int), ("unsteady", C.c_int),
                ("cfl", C.c_double), ("time_step", C.c_double), ("final_time", C.c_double),
                ("min_residue", C.c_double), ("max_iter", C.c_long),
                ("n_time_stamps", C.c_int), ("cell_type", C.c_int), ("liou_fix", C.c_int),
                ("kernel_path", C.c_int), ("tile_vertices", C.c_int),
                ("block_grad", C.c_int), ("block_fu", C.c_int),
                ("tile_fused_gradient", C.c_int), ("graph_steps", C.c_int),
                ("host_preprocess", C.c_int), ("strong_bc", C.c_int)]


class TsmMesh(C.Structure):
    _fields_ = [("n_vertex", C.c_int), ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("n_bface", C.c_int), ("bface_tag", C.POINTER(C.c_int)),
                ("bface_v", C.POINTER(C.c_int)),
                ("n_cell", C.c_int), ("cell_v", C.POINTER(C.c_int)),
                ("n_periodic", C.c_int), ("periodic_ptr", C.POINTER(C.c_int)),
                ("periodic_v", C.POINTER(C.c_int)),
                ("n_bc", C.c_int), ("bc_tag", C.POINTER(C.c_int)),
                ("bc_type", C.POINTER(C.c_int)), ("bc_state", C.POINTER(C.c_double)),
                ("bc_adiabatic", C.POINTER(C.c_int))]


@dataclass
class Params:
    """Parsed numeric/material settings (names as in param.in)."""
    gamma: float = 1.4
    gas_const: float = 1.0
    prandtl: float = 0.72
    mu_ref: float = 0.0
    T_ref: float = 1.0
    T_0: float = 0.0
    omega: float = 1.0
    model: str = "euler"
    viscosity: str = "constant"
    flux: str = "kepes_tecno_rusanov"
    reconstruct: str = "minmod"
    tvb_M: float = 0.0
    ev_alpha: float = 0.0
    ev_beta: float = 0.0
    nd_active: bool = True
    visc_active: bool = True
    heat_active: bool = True
    time_scheme: str = "ssprk3"
    time_mode: str = "unsteady"
    cfl: float = 0.5
    time_step: float = 0.0
    final_time: float = 1.0
    min_residue: float = 0.0
    max_iter: int = 1000000
    time_stamps: int = 1
    cell: str = "median"
    liou_fix: bool = False
    bc_scheme: str = "weak"        # "weak" or "strong" (boundary values imposed after each stage)
    kernel_path: str = "auto"      # "auto" (tiled where supported) or "csr"
    tile_vertices: int = 0
    block_grad: int = 0            # threads per tile, gradient kernel (0: auto)
    block_fu: int = 0              # threads per tile, flux/update kernel (0: auto)
    fused_gradient: bool = False   # gradients inside the stage kernel (two halo rings)
    graph_steps: int = 0           # steps per captured HIP graph (0: no graph)
    preprocess: str = "auto"       # "auto" (device geometry for median duals) or "host"

    def to_c(self):
        p = TsmParams()
        p.gamma, p.gas_const, p.prandtl = self.gamma, self.gas_const, self.prandtl
        p.mu_ref, p.T_ref, p.T_0, p.omega = self.mu_ref, self.T_ref, self.T_0, self.omega
        p.model = {"euler": 0, "ns": 1}[self.model]
        p.mu_model = {"constant": 0, "sutherland": 1, "power": 2}[self.viscosity]
        p.flux = FLUX[self.flux]
        p.reconstruct = RECONSTRUCT[self.reconstruct]
        p.tvb_M, p.ev_alpha, p.ev_beta = self.tvb_M, self.ev_alpha, self.ev_beta
        p.nd_active, p.visc_active, p.heat_active = int(self.nd_active), int(self.visc_active), int(self.heat_active)
        p.time_scheme = {"ssprk3": 0, "rk4": 1, "heuns": 2}[self.time_scheme]
        p.unsteady = 1 if self.time_mode == "unsteady" else 0
        p.cfl, p.time_step, p.final_time = self.cfl, self.time_step, self.final_time
        p.min_residue, p.max_iter, p.n_time_stamps = self.min_residue, self.max_iter, self.time_stamps
        p.cell_type = {"median": 0, "voronoi": 1}[self.cell]
        p.liou_fix = int(self.liou_fix)
        p.kernel_path = {"auto": 0, "csr": 1}[self.kernel_path]
        p.tile_vertices = self.tile_vertices
        p.block_grad, p.block_fu = self.block_grad, self.block_fu
        p.tile_fused_gradient = int(self.fused_gradient)
        p.graph_steps = self.graph_steps
        p.host_preprocess = {"auto": 0, "host": 1}[self.preprocess]
        p.strong_bc = {"weak": 0, "strong": 1}[self.bc_scheme]
        return p


@dataclass
class Mesh:
    """NPART = 1 part-file content (0-based ids) plus the boundary table."""
    x: np.ndarray
    y: np.ndarray
    bface_tag: np.ndarray
    bface_v: np.ndarray
    cell_v: np.ndarray
    periodic_ptr: np.ndarray
    periodic_v: np.ndarray
    # tag -> (type name, (T,u,v,p)[, adiabatic]); a noslip wall is isothermal at
    # its state's T unless adiabatic is true (slip walls always are)
    bc: dict = field(default_factory=dict)

    @property
    def n_vertex(self):
        return int(self.x.size)

    def to_c(self):
        keep = []

        def ptr(a, t):
            a = np.ascontiguousarray(a, dtype=np.float64 if t is C.c_double else np.int32)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(t))

        m = TsmMesh()
        m.n_vertex = self.n_vertex
        m.x, m.y = ptr(self.x, C.c_double), ptr(self.y, C.c_double)
        m.n_bface = int(self.bface_tag.size)
        m.bface_tag, m.bface_v = ptr(self.bface_tag, C.c_int), ptr(self.bface_v, C.c_int)
        m.n_cell = int(self.cell_v.size // 3)
        m.cell_v = ptr(self.cell_v, C.c_int)
        m.n_periodic = int(self.periodic_ptr.size - 1) if self.periodic_ptr.size else 0
        m.periodic_ptr = ptr(self.periodic_ptr if self.periodic_ptr.size else np.zeros(1), C.c_int)
        m.periodic_v = ptr(self.periodic_v if self.periodic_v.size else np.zeros(1), C.c_int)
        tags = sorted(self.bc)
        m.n_bc = len(tags)
        m.bc_tag = ptr(np.array(tags, dtype=np.int32), C.c_int)
        m.bc_type = ptr(np.array([BC[self.bc[t][0]] for t in tags], dtype=np.int32), C.c_int)
        st = np.array([self.bc[t][1] if len(self.bc[t]) > 1 else (0, 0, 0, 0) for t in tags],
                      dtype=np.float64).reshape(-1)
        m.bc_state = ptr(st if st.size else np.zeros(4), C.c_double)
        adia = [int(self.bc[t][2]) if len(self.bc[t]) > 2 else int(self.bc[t][0] == "slip")
                for t in tags]
        m.bc_adiabatic = ptr(np.array(adia if adia else [0], dtype=np.int32), C.c_int)
        return m, keep


_lib = None


def lib():
    """Loads libtsm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libtsm.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.tsm_last_error.restype = C.c_char_p
        L.tsm_create.argtypes = [C.POINTER(TsmMesh), C.POINTER(TsmParams), C.c_int,
                                 C.POINTER(C.c_void_p)]
        L.tsm_destroy.argtypes = [C.c_void_p]
        L.tsm_destroy.restype = None
        for name in ("tsm_set_state", "tsm_get_state", "tsm_residual", "tsm_gradients"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_set_sample_material.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.tsm_compute_dt.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.tsm_run.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
        L.tsm_time.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)]
        L.tsm_status.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_long),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.tsm_residual_norms.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_set_bc_state.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_set_bc_time_expr.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_void_p]
        L.tsm_forces.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_keep_stage_state.argtypes = [C.c_void_p, C.c_int]
        L.tsm_get_clock.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4
        L.tsm_bounds.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_status_all.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.tsm_set_clock.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_int, C.c_double, C.c_double]
        L.tsm_stage_gradients.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_surface.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_error_norms.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.tsm_global.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.tsm_geometry_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.tsm_geometry.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.tsm_bench.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_sobol.argtypes = [C.c_long, C.c_int, C.c_long, C.c_ulonglong, C.c_int, C.c_void_p]
        L.tsm_step.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        L.tsm_run_group.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_void_p]
        L.tsm_run_packed.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_void_p]
        L.tsm_pert.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_int,
                               C.c_void_p]
        L.tsm_moments.argtypes = [C.c_void_p] + [C.c_void_p] * 3
        L.tsm_probe_density.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 3
        L.tsm_probe_density_noalloc.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 3
        L.tsm_check_arith.argtypes = [C.c_long] + [C.c_void_p] * 3
        L.tsm_enstrophy.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.tsm_global_all.argtypes = [C.c_void_p] + [C.c_void_p] * 2
        L.tsm_mesh_jittered_sizes.argtypes = [C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 5
        L.tsm_mesh_jittered.argtypes = ([C.c_int] + [C.c_double] * 5 + [C.c_ulonglong, C.c_int]
                                        + [C.c_void_p] * 7)
        L.tsm_host_tiling.argtypes = [C.POINTER(TsmMesh), C.c_int, C.c_int, C.c_void_p]
        L.tsm_host_geometry.argtypes = [C.POINTER(TsmMesh), C.c_int] + [C.c_void_p] * 6 + [C.POINTER(C.c_int)] * 2
        L.tsm_device_geometry.argtypes = L.tsm_host_geometry.argtypes
        _lib = L
    return _lib


class TsmError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise TsmError(f"libtsm error {rc}: {lib().tsm_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(C.c_void_p)


def read_part_file(path, bc):
    """Mesh from an NPART = 1 part file (PARTITION/part_file_0.dat, layout of
    readme/README.md:155-201, read as src/read_part_mesh_file.cc:43-214 reads
    it): nodes with their global ids, periodic sets, boundary faces (tag, two
    gids, periodic flag), cells (three gids, ghost flag).  Global ids map to
    0-based vertices in file order; bc is the boundary table {tag: (type, ...)}
    as for Mesh.  Ghost cells (NPART > 1) are refused."""
    with open(path) as f:
        tok = f.read().split()
    pos = 0

    def take(n=1):
        nonlocal pos
        out = tok[pos:pos + n]
        if len(out) < n:
            raise TsmError(f"part file truncated: {path}")
        pos += n
        return out

    def expect(word):
        w = take()[0]
        if w != word:
            raise TsmError(f"part file {path}: {word} expected, found {w}")

    expect("DIMENSION")
    if int(take()[0]) != 2:
        raise TsmError("part file: DIMENSION 2 expected")
    expect("PARTITIONS")
    if int(take()[0]) != 1:
        raise TsmError("part file: one partition expected (NPART = 1)")
    expect("NODES")
    n = int(take()[0])
    expect("OF")
    if int(take()[0]) != n or n <= 0:
        raise TsmError("part file: NODES")
    gid = np.empty(n, dtype=np.int64)
    x, y = np.empty(n), np.empty(n)
    for i in range(n):
        g, xi, yi, _z, ns = take(5)
        gid[i], x[i], y[i] = int(g), float(xi), float(yi)
        take(int(ns) + 1)  # owning partitions, boundary flag
    g2l = {int(g): i for i, g in enumerate(gid)}
    expect("PERIODIC")
    ptr, mem = [0], []
    for _ in range(int(take()[0])):
        _id, cnt = take(2)
        mem += [g2l[int(v)] for v in take(int(cnt))]
        take(int(take()[0]))  # partitions of the set
        ptr.append(len(mem))
    expect("FACES")
    nb = int(take()[0])
    expect("OF")
    take()
    btag, bv = np.empty(nb, dtype=np.int32), np.empty(2 * nb, dtype=np.int32)
    for k in range(nb):
        t, a, b, _per = take(4)
        btag[k], bv[2 * k], bv[2 * k + 1] = int(t), g2l[int(a)], g2l[int(b)]
    expect("CELLS")
    nc = int(take()[0])
    expect("OF")
    take()
    cv = np.empty(3 * nc, dtype=np.int32)
    for k in range(nc):
        a, b, c, ghost = take(4)
        if int(ghost):
            raise TsmError("part file: ghost cells need NPART > 1")
        cv[3 * k:3 * k + 3] = (g2l[int(a)], g2l[int(b)], g2l[int(c)])
    if len(ptr) == 1:  # no periodic sets: empty arrays, as jittered_mesh makes them
        ptr = []
    return Mesh(x, y, btag, bv, cv, np.asarray(ptr, dtype=np.int32), np.asarray(mem, dtype=np.int32),
                bc)


def jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=0, periodic=False, bc=None):
    """Seeded jittered-triangle n x n lattice from libtsm's generator."""
    L = lib()
    s = [C.c_int() for _ in range(5)]
    _check(L.tsm_mesh_jittered_sizes(n, int(periodic), *[C.byref(v) for v in s]))
    nv, nb, nc, npr, npv = [v.value for v in s]
    x, y = np.empty(nv), np.empty(nv)
    bt, bv = np.empty(nb, np.int32), np.empty(2 * nb, np.int32)
    cv = np.empty(3 * nc, np.int32)
    pp, pv = np.empty(npr + 1 if periodic else 0, np.int32), np.empty(npv, np.int32)
    _check(L.tsm_mesh_jittered(n, box[0], box[1], box[2], box[3], jitter, seed, int(periodic),
                               _dptr(x), _dptr(y), _dptr(bt), _dptr(bv), _dptr(cv),
                               _dptr(pp) if periodic else None, _dptr(pv) if periodic else None))
    if bc is None:
        bc = {t: ("periodic",) if periodic else ("outlet",) for t in (1, 2, 3, 4)}
    return Mesh(x, y, bt, bv, cv, pp, pv, bc)


def check_arith(a, b):
    """The stage kernels' fdiv/fsqrt beside the device's IEEE / and sqrt():
    (n, 4) array of fdiv(a, b), a / b, fsqrt(a), sqrt(a)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty((a.size, 4))
    _check(lib().tsm_check_arith(a.size, _dptr(a), _dptr(b), _dptr(out)))
    return out


def host_geometry(mesh: Mesh, cell="median"):
    """Geometry of libtsm's host preprocessing (no device needed)."""
    return _geometry(lib().tsm_host_geometry, mesh, cell)


def device_geometry(mesh: Mesh, cell="median"):
    """The same arrays built on the device (median duals)."""
    return _geometry(lib().tsm_device_geometry, mesh, cell)


def _geometry(fn, mesh, cell):
    L = lib()
    m, keep = mesh.to_c()
    nf, nb = C.c_int(), C.c_int()
    _check(fn(C.byref(m), {"median": 0, "voronoi": 1}[cell], None, None, None, None, None, None,
              C.byref(nf), C.byref(nb)))
    fv = np.empty(2 * nf.value, np.int32)
    fnv = np.empty(2 * nf.value)
    bvv = np.empty(2 * nb.value, np.int32)
    bn = np.empty(2 * nb.value)
    dc, mc = np.empty(mesh.n_vertex), np.empty(mesh.n_vertex)
    _check(fn(C.byref(m), {"median": 0, "voronoi": 1}[cell], _dptr(fv), _dptr(fnv), _dptr(bvv),
              _dptr(bn), _dptr(dc), _dptr(mc), C.byref(nf), C.byref(nb)))
    return dict(face_v=fv.reshape(-1, 2), face_n=fnv.reshape(-1, 2), bface_v=bvv.reshape(-1, 2),
                bface_n=bn.reshape(-1, 2), dcarea=dc, mcarea=mc)


def device_tiling_check(mesh: Mesh, tile_vertices=128):
    """The vertex tiling built on the device against the host's build (median
    duals without periodic sets): (differing entries over all arrays, name of
    the first array that differs or "")."""
    m, keep = mesh.to_c()
    n = C.c_longlong(0)
    which = C.create_string_buffer(32)
    _check(lib().tsm_device_tiling_check(C.byref(m), int(tile_vertices), C.byref(n), which))
    return n.value, which.value.decode()


def host_tiling(mesh: Mesh, tile_vertices=128, cell="median", two_rings=False):
    """Builds and bounds-checks the tiled kernels' vertex tiling on the host."""
    m, keep = mesh.to_c()
    st = np.zeros(10, dtype=np.int32)
    _check(lib().tsm_host_tiling(C.byref(m), {"median": 0, "voronoi": 1}[cell], tile_vertices,
                                 int(two_rings), _dptr(st)))
    return dict(zip(("tiles", "periodic_tiles", "max_slots", "max_faces", "halo", "faces",
                     "two_rings", "halo2", "grad_cells", "max_grad_slots"), map(int, st)))


def read_raw(path):
    """(time, (n_vertex, 4) array of T, u, v, p) from a driver output file
    written with `format raw`."""
    with open(path, "rb") as f:
        head = f.read(64)
        if head[:8] != b"TSMRAW01":
            raise TsmError(f"{path}: not a raw solution file")
        nv, nf = np.frombuffer(head[8:24], dtype="<i8")
        t = float(np.frombuffer(head[24:32], dtype="<f8")[0])
        data = np.fromfile(f, dtype="<f8", count=int(nv * nf))
    return t, data.reshape(int(nv), int(nf))


def sobol(n, dim, seed=0, scramble=True, skip=0):
    """(n, dim) Sobol points in [0,1)^dim, Gray-code order, optionally Owen-scrambled."""
    out = np.empty((n, dim))
    _check(lib().tsm_sobol(n, dim, skip, seed, int(scramble), _dptr(out)))
    return out


def pert(ptype, x, y, p, rnd):
    """The reference's pert(type, x, y, p) with one sample's random numbers."""
    r = np.ascontiguousarray(rnd, dtype=np.float64)
    out = C.c_double()
    _check(lib().tsm_pert(ptype, x, y, p, _dptr(r), r.size, C.byref(out)))
    return out.value


def run_packed(solvers, nsteps):
    """Advances several Solvers (the levels of a multilevel run) with one
    launch per kernel spanning all of them; returns the steps each performed,
    summed over its samples.  Euler, SSP-RK3, tiled kernels, one scheme."""
    hs = (C.c_void_p * len(solvers))(*[s._h for s in solvers])
    done = np.zeros(len(solvers), dtype=np.int64)
    _check(lib().tsm_run_packed(hs, len(solvers), nsteps, _dptr(done)))
    return done


def run_group(solvers, nsteps):
    """Advances several Solvers together (one stream each); returns the steps
    each performed, summed over its samples."""
    hs = (C.c_void_p * len(solvers))(*[s._h for s in solvers])
    done = np.zeros(len(solvers), dtype=np.int64)
    _check(lib().tsm_run_group(hs, len(solvers), nsteps, _dptr(done)))
    return done


class Solver:
    """One libtsm context: the whole mesh resident on the GPU, n_samples states."""

    def __init__(self, mesh: Mesh, params: Params, n_samples: int = 1):
        L = lib()
        self.mesh, self.params, self.n_samples = mesh, params, n_samples
        m, self._keep = mesh.to_c()
        self._p = params.to_c()
        h = C.c_void_p()
        _check(L.tsm_create(C.byref(m), C.byref(self._p), n_samples, C.byref(h)))
        self._h = h
        self.nv = mesh.n_vertex
        self.n_bface = int(mesh.bface_tag.size)
        self._bface_tag = np.asarray(mesh.bface_tag, dtype=np.int32)

    def close(self):
        if getattr(self, "_h", None):
            lib().tsm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_material(self, sample, gamma, gas_const):
        _check(lib().tsm_set_sample_material(self._h, sample, gamma, gas_const))

    def set_state(self, prim, sample=0):
        a = np.ascontiguousarray(prim, dtype=np.float64).reshape(self.nv, 4)
        _check(lib().tsm_set_state(self._h, sample, _dptr(a)))

    def get_state(self, sample=0):
        a = np.empty((self.nv, 4))
        _check(lib().tsm_get_state(self._h, sample, _dptr(a)))
        return a

    def residual(self, sample=0):
        a = np.empty((self.nv, 4))
        _check(lib().tsm_residual(self._h, sample, _dptr(a)))
        return a

    def gradients(self, sample=0):
        a = np.empty((self.nv, 8))
        _check(lib().tsm_gradients(self._h, sample, _dptr(a)))
        return a

    def compute_dt(self, sample=0):
        d = C.c_double()
        _check(lib().tsm_compute_dt(self._h, sample, C.byref(d)))
        return d.value

    def run(self, nsteps, want_log=False):
        done = np.zeros(self.n_samples, dtype=np.int64)
        log = np.zeros((self.n_samples, nsteps, 4)) if want_log else None
        _check(lib().tsm_run(self._h, nsteps, _dptr(done), _dptr(log) if want_log else None))
        return (done, log[0] if self.n_samples == 1 else log) if want_log else done

    def step(self, t_target=0.0):
        """One time step of every running sample (optionally landing on t_target);
        returns each sample's dt and residual norm."""
        dt, res = np.zeros(self.n_samples), np.zeros(self.n_samples)
        _check(lib().tsm_step(self._h, t_target, _dptr(dt), _dptr(res)))
        return dt, res

    def status(self, sample=0):
        v = np.zeros(3)
        it, cnt, dn = C.c_long(), C.c_int(), C.c_int()
        _check(lib().tsm_status(self._h, sample, _dptr(v), C.byref(it), C.byref(cnt), C.byref(dn)))
        return dict(t=v[0], dt=v[1], res=v[2], iter=it.value, counter=cnt.value, done=bool(dn.value))

    def status_all(self):
        """Every sample's loop state in one transfer: dict of arrays t, dt, res,
        iter, counter, done and norms (n_samples, 4)."""
        n = self.n_samples
        v, rn = np.zeros((n, 3)), np.zeros((n, 4))
        it, cnt, dn = np.zeros(n, np.int64), np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().tsm_status_all(self._h, _dptr(v), it.ctypes.data_as(C.c_void_p),
                                    cnt.ctypes.data_as(C.c_void_p), dn.ctypes.data_as(C.c_void_p),
                                    _dptr(rn)))
        return dict(t=v[:, 0], dt=v[:, 1], res=v[:, 2], iter=it, counter=cnt, done=dn.astype(bool),
                    norms=rn)

    def set_bc_state(self, values, sample=0):
        """Boundary values of one sample per bface end point: values (n_bface,
        2, 4) = (T,u,v,p) at the vertices bface_v[b, 0] and bface_v[b, 1]."""
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != 8 * self.n_bface:
            raise TsmError(f"expected {8 * self.n_bface} boundary values, got {a.size}")
        _check(lib().tsm_set_bc_state(self._h, sample, _dptr(a)))

    def set_bc_time_expr(self, groups, exprs, constants=None):
        """Boundary values that depend on t (Heun runs): groups (n_bface,)
        names each bface's group (-1: none); exprs[g] = (T, u, v, p)
        expressions of x, y and t for group g, None keeping a field's value;
        constants = {name: value}.  An empty exprs removes them."""
        g = np.ascontiguousarray(groups, dtype=np.int32).reshape(-1)
        if g.size != self.n_bface:
            raise TsmError(f"expected {self.n_bface} bface groups, got {g.size}")
        flat = [e for four in exprs for e in four]
        if len(flat) != 4 * len(exprs):
            raise TsmError("each group needs four expressions (T, u, v, p)")
        ex = (C.c_char_p * max(1, len(flat)))(*[None if e is None else e.encode() for e in flat])
        constants = dict(constants or {})
        names = (C.c_char_p * max(1, len(constants)))(*[k.encode() for k in constants])
        vals = np.array(list(constants.values()) or [0.0], dtype=np.float64)
        _check(lib().tsm_set_bc_time_expr(self._h, g.ctypes.data_as(C.c_void_p), len(exprs),
                                          C.cast(ex, C.c_void_p), len(constants),
                                          C.cast(names, C.c_void_p), _dptr(vals)))

    def forces(self, tags, sample=0, force_index=0):
        """(pressure x, y, viscous x, y) force on the boundary faces with these tags."""
        t = np.ascontiguousarray(tags, dtype=np.int32)
        out = np.zeros(4)
        _check(lib().tsm_forces(self._h, sample, int(force_index), t.ctypes.data_as(C.c_void_p),
                                t.size, _dptr(out)))
        return out

    def bounds(self, sample=0):
        """(10,) minima and maxima of T, p, u, v and w (zero) of the current state."""
        out = np.zeros(10)
        _check(lib().tsm_bounds(self._h, sample, _dptr(out)))
        return out

    def clock(self, sample=0):
        """(iter, counter, t, res0): the loop clock a restart file carries."""
        it, cn, t, r0 = C.c_long(), C.c_int(), C.c_double(), C.c_double()
        _check(lib().tsm_get_clock(self._h, sample, C.byref(it), C.byref(cn), C.byref(t),
                                   C.byref(r0)))
        return it.value, cn.value, t.value, r0.value

    def set_clock(self, it, counter, t, res0, sample=0):
        """Continue from a restart: set after set_state."""
        _check(lib().tsm_set_clock(self._h, sample, int(it), int(counter), float(t), float(res0)))

    def keep_stage_state(self, on=True):
        """Keep each step's last-stage input state, whose gradients the
        reference's outputs after a step use (stage_gradients, surface)."""
        _check(lib().tsm_keep_stage_state(self._h, int(bool(on))))

    def stage_gradients(self, sample=0):
        """(n_vertex, 8) gradients of the kept last-stage input state."""
        a = np.empty((self.nv, 8))
        _check(lib().tsm_stage_gradients(self._h, sample, _dptr(a)))
        return a

    def surface(self, tags, sample=0, stage=True):
        """(n, 2) skin friction and heat flux of the boundary faces with these
        tags, in bface order; gradients of the kept last-stage state (stage)
        or of the current state."""
        t = np.ascontiguousarray(tags, dtype=np.int32)
        out = np.zeros((int(np.isin(self._bface_tag, t).sum()), 2))
        _check(lib().tsm_surface(self._h, sample, int(bool(stage)), t.ctypes.data_as(C.c_void_p),
                                 t.size, _dptr(out)))
        return out

    def residual_norms(self, sample=0):
        """Last step's residual norms (mass, x/y momentum, energy) before
        normalisation, as the steady-mode residual log prints them."""
        out = np.zeros(4)
        _check(lib().tsm_residual_norms(self._h, sample, _dptr(out)))
        return out

    def time(self, sample=0):
        t, it = C.c_double(), C.c_long()
        _check(lib().tsm_time(self._h, sample, C.byref(t), C.byref(it)))
        return t.value, it.value

    def error_norms(self, exact_prim, sample=0):
        a = np.ascontiguousarray(exact_prim, dtype=np.float64).reshape(self.nv, 4)
        out = np.empty(24)
        _check(lib().tsm_error_norms(self._h, sample, _dptr(a), _dptr(out)))
        return out.reshape(8, 3)

    def global_integrals(self, sample=0):
        k, e = C.c_double(), C.c_double()
        _check(lib().tsm_global(self._h, sample, C.byref(k), C.byref(e)))
        return k.value, e.value

    def enstrophy(self, sample=0):
        """1/2 sum omega^2 dcarea w of the current state."""
        z = C.c_double()
        _check(lib().tsm_enstrophy(self._h, sample, C.byref(z)))
        return z.value

    def moments(self):
        """Per-vertex (mean, M2, S2) of (T, u, v, p, rho) over all samples, each (n_vertex, 5)."""
        m, q, z = (np.empty((self.nv, 5)) for _ in range(3))
        _check(lib().tsm_moments(self._h, _dptr(m), _dptr(q), _dptr(z)))
        return m, q, z

    def probe_density(self, px, py):
        """Density of every sample at the probe points: (n_samples, n_probe)."""
        px = np.ascontiguousarray(px, dtype=np.float64)
        py = np.ascontiguousarray(py, dtype=np.float64)
        out = np.empty((self.n_samples, px.size))
        _check(lib().tsm_probe_density(self._h, px.size, _dptr(px), _dptr(py), _dptr(out)))
        return out

    def probe_density_noalloc(self, px, py):
        """probe_density without device allocation (the library's fixed scratch
        array; calls that overlap on one GPU race)."""
        px = np.ascontiguousarray(px, dtype=np.float64)
        py = np.ascontiguousarray(py, dtype=np.float64)
        out = np.empty((self.n_samples, px.size))
        _check(lib().tsm_probe_density_noalloc(self._h, px.size, _dptr(px), _dptr(py), _dptr(out)))
        return out

    def global_all(self):
        """Kinetic energy and entropy integrals of every sample: two (n_samples,) arrays."""
        k, e = np.empty(self.n_samples), np.empty(self.n_samples)
        _check(lib().tsm_global_all(self._h, _dptr(k), _dptr(e)))
        return k, e

    def geometry(self):
        nf, nb = C.c_int(), C.c_int()
        _check(lib().tsm_geometry_sizes(self._h, C.byref(nf), C.byref(nb)))
        fv = np.empty(2 * nf.value, np.int32)
        fn = np.empty(2 * nf.value)
        dc, mc = np.empty(self.nv), np.empty(self.nv)
        _check(lib().tsm_geometry(self._h, _dptr(fv), _dptr(fn), _dptr(dc), _dptr(mc)))
        return dict(face_v=fv.reshape(-1, 2), face_n=fn.reshape(-1, 2), dcarea=dc, mcarea=mc)

    def bench(self, nsteps):
        us = np.zeros(5)
        _check(lib().tsm_bench(self._h, nsteps, _dptr(us)))
        return dict(dt_us=us[0], gradient_us=us[1], flux_us=us[2], update_us=us[3], step_us=us[4])
