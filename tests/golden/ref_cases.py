"""Cases run by the reference solver to produce the fixtures ref_<name>.npz.

Each case is one param.in (the reference's own grammar, src/parameter.cc) on a
part file of this build's seeded jittered mesh, plus the numpy evaluation of
the same initial-state and exact-solution expressions, so that the tests can
start this build's solver (and the oracle) from the state the reference
starts from.  make_ref_fixtures.py writes the fixtures; the tests read them.

What a fixture holds (make_ref_fixtures.py):
  log    residual.dat of the run, rows (iter, dt, t, res_total) at the 12
         digits the reference prints (src/fv.cc:1959-1968);
  state  the restart file written after step `steps` (T, u, v, p per vertex at
         16 significant digits, src/writer.cc:382-409; w = 0 dropped);
  clock  the loop clock that restart file carries (iter, counter, t, res0);
  error  error_norms.dat, 8 variables x (L1, L2, Linf), when the case asks for
         it (src/fv.cc:1712-1795, written at %.8e);
  glob   global.dat rows (t, KE, entropy) at 6 digits, when asked.
"""
from __future__ import annotations

import numpy as np

import tensum_amd as T
from cases import riemann, riemann_mesh, taylor_green, tgv_mesh, vortex, vortex_mesh

VORTEX_IC = {
    "temperature": "1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x,2)-pow(y,2)))",
    "xvelocity": "1 - (beta/(2*pi))*(y)*exp(0.5*(1-pow(x,2)-pow(y,2)))",
    "yvelocity": "1 + (beta/(2*pi))*(x)*exp(0.5*(1-pow(x,2)-pow(y,2)))",
    "pressure": "pow(1.0 - (gam-1)*(pow(beta,2)/(8*gam*pi*pi))*exp((1-pow(x,2)-pow(y,2))),gam/(gam-1))",
}
# the exact solution at time t without a periodic wrap (as the expressions read)
VORTEX_EXACT = {k: v.replace("pow(x,2)", "pow(x-t,2)").replace("pow(y,2)", "pow(y-t,2)")
                .replace("(y)*", "(y-t)*").replace("(x)*", "(x-t)*") for k, v in VORTEX_IC.items()}

# quadrant states split at (0.8, 0.8): T = p / (R rho)
RIEMANN_IC = {
    "temperature": "if(x>=0.8, if(y>=0.8, 1.5/1.5, 0.3/0.5323), if(y>=0.8, 0.3/0.5323, 0.029/0.138))",
    "xvelocity": "if(x>=0.8, 0.0, 1.206)",
    "yvelocity": "if(y>=0.8, 0.0, 1.206)",
    "pressure": "if(x>=0.8, if(y>=0.8, 1.5, 0.3), if(y>=0.8, 0.3, 0.029))",
}

TGV_IC = {
    "temperature": "p0 + 0.25*(cos(2*x)+cos(2*y))",
    "xvelocity": "sin(x)*cos(y)",
    "yvelocity": "-cos(x)*sin(y)",
    "pressure": "p0 + 0.25*(cos(2*x)+cos(2*y))",
}

# a smooth state near the boundary values of the mixed-boundary case
MIXED_STATE = (1.0, 0.3, 0.1, 1.2)
MIXED_IC = {
    "temperature": "1.0*(1+0.1*sin(2*pi*x)*sin(2*pi*y))",
    "xvelocity": "0.3",
    "yvelocity": "0.1",
    "pressure": "1.2*(1+0.05*cos(2*pi*x))",
}


def _mixed_state(x, y):
    w = np.tile(np.array(MIXED_STATE), (x.size, 1))
    w[:, 0] *= 1 + 0.1 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
    w[:, 3] *= 1 + 0.05 * np.cos(2 * np.pi * x)
    return w


def _mixed_mesh(n, kinds):
    st = MIXED_STATE
    bc = {}
    for tag, kind in zip((1, 2, 3, 4), kinds):
        if kind in ("inlet", "farfield", "pressure"):
            bc[tag] = (kind, st)
        elif kind == "noslip":  # adiabatic wall at rest
            bc[tag] = (kind, (st[0], 0.0, 0.0, st[3]), True)
        elif kind == "noslip_T":  # isothermal wall at rest, T = 1.1
            bc[tag] = ("noslip", (1.1, 0.0, 0.0, st[3]))
        else:
            bc[tag] = (kind,)
    return T.jittered_mesh(n, seed=5, bc=bc)


def _vortex_case(flux, rec, n=17, steps=60, cfl=0.8, **kw):
    return dict(mesh=lambda: vortex_mesh(n), ic=VORTEX_IC, consts={"gam": 1.4, "beta": 5},
                state=lambda m: vortex(m.x, m.y, box=None), bc={(1, 2, 3, 4): ("periodic", None)},
                params=dict(flux=flux, reconstruct=rec, cfl=cfl, **kw), steps=steps, final_time=None)


def _riemann_case(flux, rec, n=25, steps=100, **kw):
    return dict(mesh=lambda: riemann_mesh(n), ic=RIEMANN_IC, consts={},
                state=lambda m: riemann(m.x, m.y), bc={(1, 2, 3, 4): ("outlet", None)},
                params=dict(flux=flux, reconstruct=rec, cfl=0.5, **kw), steps=steps, final_time=None)


# name -> settings.  `steps`: the restart after that step is kept; the run
# goes on to final_time (unsteady runs ignore max_iter, src/fv.cc:2265-2275).
CASES = {
    # the workload of the throughput benchmark, on a small mesh
    "riemann33": dict(
        mesh=lambda: riemann_mesh(33), ic=RIEMANN_IC, consts={},
        state=lambda m: riemann(m.x, m.y), bc={(1, 2, 3, 4): ("outlet", None)},
        params=dict(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5),
        steps=334, final_time=None),
    # the smooth periodic case of the parity bar, with its error norms
    "vortex33": dict(
        mesh=lambda: vortex_mesh(33), ic=VORTEX_IC, consts={"gam": 1.4, "beta": 5},
        state=lambda m: vortex(m.x, m.y, box=None), bc={(1, 2, 3, 4): ("periodic", None)},
        params=dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8),
        steps=334, final_time=None, exact=VORTEX_EXACT,
        exact_state=lambda m, t: vortex(m.x, m.y, t, box=None)),
    # every boundary type with a state (slip, outlet, inlet, farfield)
    "mixed23": dict(
        mesh=lambda: _mixed_mesh(23, ("slip", "outlet", "inlet", "farfield")), ic=MIXED_IC,
        consts={}, state=lambda m: _mixed_state(m.x, m.y),
        bc={(1,): ("slip", None), (2,): ("outlet", None), (3,): ("inlet", MIXED_STATE),
            (4,): ("farfield", MIXED_STATE)},
        params=dict(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.4),
        steps=100, final_time=None),
    # pressure outlet opposite an inlet, Roe flux, van Albada
    "pressure23": dict(
        mesh=lambda: _mixed_mesh(23, ("slip", "pressure", "slip", "inlet")), ic=MIXED_IC,
        consts={}, state=lambda m: _mixed_state(m.x, m.y),
        bc={(1, 3): ("slip", None), (2,): ("pressure", MIXED_STATE), (4,): ("inlet", MIXED_STATE)},
        params=dict(flux="kepes_tecno_roe", reconstruct="van_albada", cfl=0.4),
        steps=100, final_time=None),
    # the Liou fix with eigenvalue modification on the Riemann problem
    "liou25": dict(
        mesh=lambda: riemann_mesh(25), ic=RIEMANN_IC, consts={},
        state=lambda m: riemann(m.x, m.y), bc={(1, 2, 3, 4): ("outlet", None)},
        params=dict(flux="kepes_tecno_roe", reconstruct="tvb_minmod", tvb_M=10.0, cfl=0.5,
                    liou_fix=True, ev_alpha=0.3, ev_beta=0.1),
        steps=100, final_time=None),
    # the reference's isentropic-vortex example settings: RK4, KEPEC, second
    # order, Voronoi duals
    "vortex17_rk4_voronoi": dict(
        mesh=lambda: vortex_mesh(17), ic=VORTEX_IC, consts={"gam": 1.4, "beta": 5},
        state=lambda m: vortex(m.x, m.y, box=None), bc={(1, 2, 3, 4): ("periodic", None)},
        params=dict(flux="kepec", reconstruct="second", cfl=0.4, time_scheme="rk4",
                    cell="voronoi"),
        steps=60, final_time=None),
    # Navier-Stokes Taylor-Green vortex (Mach 0.1, Re 1600, Pr 0.72), KEPEC
    "tgv33_ns": dict(
        mesh=lambda: tgv_mesh(33), ic=TGV_IC, consts={"p0": 1.0 / (1.4 * 0.1 ** 2)},
        state=lambda m: taylor_green(m.x, m.y), bc={(1, 2, 3, 4): ("periodic", None)},
        params=dict(model="ns", flux="kepec", reconstruct="first", cfl=0.5, mu_ref=1.0 / 1600,
                    prandtl=0.72),
        steps=200, final_time=None, save_global=True),
    # the other two-point fluxes and every reconstruction of the TeCNO fluxes
    "kep_vortex17": _vortex_case("kep", "first"),
    "roe_ec_vortex17": _vortex_case("roe_ec", "first"),
    "avg_vortex17": _vortex_case("simple_avg", "first", cfl=0.4),
    "albada_vortex25": _vortex_case("kepes_tecno_roe", "van_albada", n=25, steps=100),
    "second_vortex25": _vortex_case("kepes_tecno_rusanov", "second", n=25, steps=100),
    "eno2_riemann25": _riemann_case("kepes_tecno_rusanov", "eno2"),
    "first_riemann25": _riemann_case("kepes_tecno_roe", "first"),
    # Navier-Stokes wall cells: adiabatic noslip floor, isothermal noslip lid,
    # inlet and outlet; KEPES-TeCNO Roe with minmod and the viscous fluxes
    "walls_ns23": dict(
        mesh=lambda: _mixed_mesh(23, ("noslip", "outlet", "noslip_T", "inlet")), ic=MIXED_IC,
        consts={}, state=lambda m: _mixed_state(m.x, m.y),
        bc={(1,): ("noslip", None), (2,): ("outlet", None), (3,): ("noslip_T", 1.1),
            (4,): ("inlet", MIXED_STATE)},
        params=dict(model="ns", flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.4,
                    mu_ref=0.01, prandtl=0.72),
        steps=100, final_time=None, save_global=True),
    # the plain-method workload of the error-norm bar: 129^2 vortex to t = 10
    # (one period of the box): the exact state equals the initial one
    "vortex129_t10": dict(
        mesh=lambda: vortex_mesh(129), ic=VORTEX_IC, consts={"gam": 1.4, "beta": 5},
        state=lambda m: vortex(m.x, m.y, box=None), bc={(1, 2, 3, 4): ("periodic", None)},
        params=dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8),
        steps=None, final_time=10.0, exact=VORTEX_IC, save_global=True),
}


def params(case, final_time):
    return T.Params(final_time=final_time, time_stamps=1, **CASES[case]["params"])


def periodic_tags(case):
    tags = []
    for k, (kind, _) in CASES[case]["bc"].items():
        if kind == "periodic":
            tags += list(k)
    return tuple(tags)


def _block(name, exprs, indent="   "):
    out = [f"{name}", "{"]
    for k in ("temperature", "xvelocity", "yvelocity"):
        out.append(f"{indent}{k:12s} {exprs[k]}")
    out.append(f"{indent}{'zvelocity':12s} 0.0")
    out.append(f"{indent}{'pressure':12s} {exprs['pressure']}")
    out.append("}")
    return out


def param_text(case, final_time, restart_frequency):
    c = CASES[case]
    p = dict(c["params"])
    ns = p.get("model") == "ns"
    tvb = p.get("tvb_M", 0.0)
    lines = ["grid", "{", f"   cell      {p.get('cell', 'median')}", "}", "numeric", "{",
             "   time_mode    unsteady", f"   time_scheme  {p.get('time_scheme', 'ssprk3')}",
             "   time_step    0.0", f"   cfl          {p['cfl']!r}", "   max_iter     1000000",
             f"   final_time   {final_time!r}", "   min_residue  0.0", "   reconstruct", "   {",
             f"      method   {p['reconstruct']}"]
    if p["reconstruct"] == "tvb_minmod":
        lines.append(f"      M_value  {tvb!r}")
    lines += ["   }", "   bc_scheme    weak", f"   liou_fix     {'yes' if p.get('liou_fix') else 'no'}",
              "   sample_list", "   {", "      groups", "      {", "      }", "      free_list",
              "      {", "      }", "   }", "   rnd_file_loc    rnd_nos.dat", "   rnd_per_sample  0",
              "}", "material", "{", "   gamma       1.4", "   gas_const   1",
              f"   model       {'ns' if ns else 'euler'}", "   viscosity", "   {",
              "      model    constant", f"      mu_ref   {p.get('mu_ref', 0.0)!r}", "   }",
              f"   prandtl     {p.get('prandtl', 0.72)!r}", f"   flux        {p['flux']}",
              "   flux_par", "   {", f"      alpha      {p.get('ev_alpha', 0.0)!r}",
              f"      beta       {p.get('ev_beta', 0.0)!r}", "   }", "   balance", "   {",
              "       switch_active  no", "       ND             yes", "       HEAT           yes",
              "       VISC           yes", "   }", "}", "constants", "{"]
    for k, v in c["consts"].items():
        lines.append(f"   {k}  {v!r}")
    lines.append("}")
    lines += _block("initial_condition", c["ic"])
    lines += ["boundary", "{"]
    for tags, (kind, st) in c["bc"].items():
        lines.append("   " + " ".join(str(t) for t in tags))
        lines += ["   {", f"      type         {kind}"]
        if kind in ("inlet", "farfield"):
            for k, v in zip(("temperature", "xvelocity", "yvelocity"), st[:3]):
                lines.append(f"      {k:12s} {v!r}")
            lines.append(f"      {'zvelocity':12s} 0.0")
            lines.append(f"      {'pressure':12s} {st[3]!r}")
        elif kind == "pressure":
            lines.append(f"      pressure     {st[3]!r}")
        elif kind == "noslip":
            lines += ["      xvelocity    0.0", "      yvelocity    0.0", "      zvelocity    0.0"]
        elif kind == "noslip_T":
            lines[-1] = "      type         noslip"
            lines += [f"      temperature  {st!r}", "      xvelocity    0.0",
                      "      yvelocity    0.0", "      zvelocity    0.0"]
        lines.append("   }")
    lines.append("}")
    if c.get("exact"):
        ex = _block("exact_soln", c["exact"])
        ex.insert(2, "   available    yes")
        lines += ex
    else:
        lines += ["exact_soln", "{", "   available    no", "}"]
    lines += ["integrals", "{", "}", "output", "{", "   format                vtk",
              "   frequency             1000000", "   time_stamps           1",
              "   use_online_stat_tool  no", "   variables", "   {", "   }", "   surfaces", "   {",
              "   }", "   restart               yes", f"   restart_frequency     {restart_frequency}",
              "   log_output            yes", "   soln_output           no",
              f"   save_global           {'yes' if c.get('save_global') else 'no'}",
              f"   find_error            {'yes' if c.get('exact') else 'no'}", "}"]
    return "\n".join(lines) + "\n"
