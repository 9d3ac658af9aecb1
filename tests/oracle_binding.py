"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")


class OrcParams(C.Structure):
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
                ("strong_bc", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(ORACLE_SO)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p, C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_error.restype = C.c_char_p
        L.orc_n_face.argtypes = [C.c_void_p]
        L.orc_n_bface.argtypes = [C.c_void_p]
        L.orc_faces.argtypes = [C.c_void_p] * 4
        L.orc_bfaces.argtypes = [C.c_void_p] * 4
        L.orc_areas.argtypes = [C.c_void_p] * 4
        L.orc_set_state.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get_state.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_gradients.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_residual.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_compute_dt.restype = C.c_double
        L.orc_compute_dt.argtypes = [C.c_void_p]
        L.orc_local_dt.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_bc_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_bc_time_fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_forces.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_stage_gradients.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get_clock.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.orc_set_clock.argtypes = [C.c_void_p, C.c_long, C.c_int, C.c_double, C.c_double]
        L.orc_surface.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_run.restype = C.c_long
        L.orc_run.argtypes = [C.c_void_p, C.c_long, C.c_void_p]
        L.orc_time.restype = C.c_double
        L.orc_time.argtypes = [C.c_void_p]
        L.orc_error_norms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_global.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_enstrophy.restype = C.c_double
        L.orc_enstrophy.argtypes = [C.c_void_p]
        L.orc_ssw.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_logavg.restype = C.c_double
        L.orc_logavg.argtypes = [C.c_double, C.c_double]
        L.orc_num_flux.argtypes = [C.POINTER(OrcParams)] + [C.c_void_p] * 7 + [C.c_double, C.c_void_p]
        L.orc_viscous_flux.argtypes = [C.POINTER(OrcParams)] + [C.c_void_p] * 3 + [C.c_int, C.c_int,
                                                                                   C.c_void_p]
        L.orc_limited_slope.restype = C.c_double
        L.orc_limited_slope.argtypes = [C.POINTER(OrcParams), C.c_double, C.c_double, C.c_double]
        L.orc_reconstruct.argtypes = [C.POINTER(OrcParams), C.c_double] + [C.c_void_p] * 6
        L.orc_moments.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def params(p):
    """OrcParams from a tensum_amd.Params (same field layout as tsm_params)."""
    c = p.to_c()
    o = OrcParams()
    for name, _ in OrcParams._fields_:
        setattr(o, name, getattr(c, name))
    return o


_BT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.POINTER(C.c_double))


class Oracle:
    def __init__(self, mesh, prm):
        L = lib()
        self._mc, self._keep = mesh.to_c()
        self._prm = params(prm)
        h = L.orc_create(C.byref(self._mc), C.byref(self._prm))
        if not h:
            raise RuntimeError("oracle: " + L.orc_error().decode())
        self._h = C.c_void_p(h)
        self.nv = mesh.n_vertex
        self._bface_v = np.asarray(mesh.bface_v, dtype=np.int32)
        self._bface_tag = np.asarray(mesh.bface_tag, dtype=np.int32)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def geometry(self):
        L = lib()
        nf, nb = L.orc_n_face(self._h), L.orc_n_bface(self._h)
        fv = np.empty(2 * nf, np.int32); fn = np.empty(2 * nf); ft = np.empty(nf, np.int32)
        L.orc_faces(self._h, _p(fv), _p(fn), _p(ft))
        bv = np.empty(2 * nb, np.int32); bn = np.empty(2 * nb); lc = np.empty(nb, np.int32)
        L.orc_bfaces(self._h, _p(bv), _p(bn), _p(lc))
        dc, mc = np.empty(self.nv), np.empty(self.nv)
        ca = np.empty(self._mc.n_cell)
        L.orc_areas(self._h, _p(dc), _p(mc), _p(ca))
        return dict(face_v=fv.reshape(-1, 2), face_n=fn.reshape(-1, 2), face_type=ft,
                    bface_v=bv.reshape(-1, 2), bface_n=bn.reshape(-1, 2), lcell=lc,
                    dcarea=dc, mcarea=mc, cell_area=ca)

    def set_state(self, prim):
        a = np.ascontiguousarray(prim, dtype=np.float64).reshape(self.nv, 4)
        lib().orc_set_state(self._h, _p(a))

    def get_state(self):
        a = np.empty((self.nv, 4))
        lib().orc_get_state(self._h, _p(a))
        return a

    def gradients(self):
        a = np.empty((self.nv, 8))
        lib().orc_gradients(self._h, _p(a))
        return a

    def residual(self):
        a = np.empty((self.nv, 4))
        lib().orc_residual(self._h, _p(a))
        return a

    def compute_dt(self):
        return lib().orc_compute_dt(self._h)

    def set_bc_state(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        bv = np.ascontiguousarray(self._bface_v, dtype=np.int32)
        if lib().orc_set_bc_state(self._h, _p(bv), _p(a)) != 0:
            raise RuntimeError("oracle: " + lib().orc_error().decode())

    def set_bc_time_fn(self, fn):
        """fn(t, values) updates values (n_bface, 2, 4) = (T,u,v,p) per bface
        end point in place at each Heun stage time t; None removes it."""
        bv = np.ascontiguousarray(self._bface_v, dtype=np.int32)
        self._bt_bv = bv
        if fn is None:
            self._bt_cb = None
            cb = None
        else:
            nb = bv.size // 2

            def call(_user, t, ptr):
                a = np.ctypeslib.as_array(ptr, shape=(nb, 2, 4))
                fn(t, a)
            self._bt_cb = _BT_FN(call)
            cb = C.cast(self._bt_cb, C.c_void_p)
        if lib().orc_set_bc_time_fn(self._h, _p(bv), cb, None) != 0:
            raise RuntimeError("oracle: " + lib().orc_error().decode())

    def forces(self, tags, force_index=0):
        t = np.ascontiguousarray(tags, dtype=np.int32)
        out = np.empty(4)
        lib().orc_forces(self._h, int(force_index), _p(t), t.size, _p(out))
        return out

    def clock(self):
        """(iter, counter, t, res0): the loop clock a restart file carries."""
        it, cn, t, r0 = C.c_long(), C.c_int(), C.c_double(), C.c_double()
        lib().orc_get_clock(self._h, C.byref(it), C.byref(cn), C.byref(t), C.byref(r0))
        return it.value, cn.value, t.value, r0.value

    def set_clock(self, it, counter, t, res0):
        lib().orc_set_clock(self._h, int(it), int(counter), float(t), float(res0))

    def stage_gradients(self):
        a = np.empty((self.nv, 8))
        lib().orc_stage_gradients(self._h, _p(a))
        return a

    def surface(self, tags):
        """(n, 2): skin friction and heat flux of each bface whose tag is
        listed, in the reference's bface order."""
        t = np.ascontiguousarray(tags, dtype=np.int32)
        out = np.empty((int(np.isin(self._bface_tag, t).sum()), 2))
        lib().orc_surface(self._h, _p(t), t.size, _p(out))
        return out

    def local_dt(self):
        a = np.empty(self.nv)
        lib().orc_local_dt(self._h, _p(a))
        return a

    def run(self, nsteps, want_log=False):
        log = np.zeros((max(nsteps, 1), 4))
        n = lib().orc_run(self._h, nsteps, _p(log))
        if n < 0:
            raise RuntimeError("oracle: " + lib().orc_error().decode())
        return (n, log[:n]) if want_log else n

    def time(self):
        return lib().orc_time(self._h)

    def error_norms(self, exact):
        a = np.ascontiguousarray(exact, dtype=np.float64).reshape(self.nv, 4)
        out = np.empty(24)
        lib().orc_error_norms(self._h, _p(a), _p(out))
        return out.reshape(8, 3)

    def ssw(self):
        out = np.zeros(self.nv, dtype=np.int32)
        lib().orc_ssw(self._h, out.ctypes.data_as(C.c_void_p))
        return out

    def enstrophy(self):
        return lib().orc_enstrophy(self._h)

    def global_integrals(self):
        k, e = C.c_double(), C.c_double()
        lib().orc_global(self._h, C.byref(k), C.byref(e))
        return k.value, e.value


def num_flux(prm, xl, xr, Wl, Wr, dEl, dEr, n, ssw=0.0):
    P = params(prm)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xl, xr, Wl, Wr, dEl, dEr, n)]
    F = np.empty(4)
    lib().orc_num_flux(C.byref(P), *[_p(a) for a in arrs], float(ssw), _p(F))
    return F


def viscous_flux(prm, ent, dE, n, wall=False, adiabatic=False):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (ent, dE, n)]
    F = np.empty(4)
    lib().orc_viscous_flux(C.byref(params(prm)), *[_p(a) for a in arrs], int(wall), int(adiabatic),
                           _p(F))
    return F


def logavg(a, b):
    return lib().orc_logavg(float(a), float(b))


def limited_slope(prm, ul, ur, h=0.0):
    return lib().orc_limited_slope(C.byref(params(prm)), float(ul), float(ur), float(h))


def reconstruct(prm, h, zl, zr, gl, gr):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (zl, zr, gl, gr)]
    rl, rr = np.empty(4), np.empty(4)
    lib().orc_reconstruct(C.byref(params(prm)), float(h), *[_p(a) for a in arrs], _p(rl), _p(rr))
    return rl, rr


def moments(prims, gammas, gas_consts):
    """(mean, M2, S2) of (T,u,v,p,rho) over the sample states prims (S, nv, 4)."""
    a = np.ascontiguousarray(prims, dtype=np.float64)
    S, nv = a.shape[0], a.shape[1]
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    r = np.ascontiguousarray(gas_consts, dtype=np.float64)
    m, q, z = (np.empty((nv, 5)) for _ in range(3))
    lib().orc_moments(nv, S, _p(a), _p(g), _p(r), _p(m), _p(q), _p(z))
    return m, q, z
