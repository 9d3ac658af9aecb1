"""GPU parity: libtsm (HIP, gfx950) against the CPU oracle.

Every test runs the product through the C ABI (tensum_amd -> libtsm.so) and
compares with the oracle's restatement of the reference on the same seeded
inputs.  Tolerances: geometry and host preprocessing bit-exact; per-call
results (gradients, residuals, dt) to rel-L2 1e-12 (GPU log differs from glibc
log by an ulp); whole runs of 334 SSP-RK3 steps (1002 stages) to rel-L2 1e-10
in conserved variables, the bar of the workload description.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import (conserved, rel_l2, riemann, riemann_mesh, taylor_green, tgv_mesh, vortex,
                   vortex_mesh)

pytestmark = pytest.mark.gpu

FLUXES = ["kepes_tecno_roe", "kepes_tecno_rusanov", "kepec", "roe_ec", "kep", "simple_avg"]
RECS = ["first", "second", "second_muscl", "van_albada", "minmod", "tvb_minmod", "eno2"]


def _pair(mesh, prm, state, n_samples=1):
    s = T.Solver(mesh, prm, n_samples)
    o = O.Oracle(mesh, prm)
    for k in range(n_samples):
        s.set_state(state, k)
    o.set_state(state)
    return s, o


def _perturbed_vortex(mesh, seed=3):
    rng = np.random.default_rng(seed)
    w = vortex(mesh.x, mesh.y)
    w[:, 0] *= 1 + 0.05 * rng.standard_normal(mesh.n_vertex)
    w[:, 3] *= 1 + 0.05 * rng.standard_normal(mesh.n_vertex)
    w[:, 1:3] += 0.05 * rng.standard_normal((mesh.n_vertex, 2))
    # keep periodic duplicates identical, as a periodic state must be
    for a, b in zip(mesh.periodic_ptr[:-1], mesh.periodic_ptr[1:]):
        w[mesh.periodic_v[a:b]] = w[mesh.periodic_v[a]]
    return w


def test_geometry_bitwise():
    m = riemann_mesh(33)
    s, o = _pair(m, T.Params(), riemann(m.x, m.y))
    g, og = s.geometry(), o.geometry()
    assert np.array_equal(g["face_v"], og["face_v"])
    assert np.array_equal(g["face_n"], og["face_n"])
    assert np.array_equal(g["dcarea"], og["dcarea"])


@pytest.mark.parametrize("periodic", [True, False])
def test_gradients(periodic):
    m = vortex_mesh(33) if periodic else riemann_mesh(33)
    w = _perturbed_vortex(m) if periodic else riemann(m.x, m.y)
    s, o = _pair(m, T.Params(flux="kepes_tecno_roe", reconstruct="minmod"), w)
    assert rel_l2(s.gradients(), o.gradients()) < 1e-12


@pytest.mark.parametrize("flux", FLUXES)
def test_residual_each_flux(flux):
    m = vortex_mesh(33)
    w = _perturbed_vortex(m)
    s, o = _pair(m, T.Params(flux=flux, reconstruct="minmod"), w)
    assert rel_l2(s.residual(), o.residual()) < 1e-12


@pytest.mark.parametrize("rec", RECS)
@pytest.mark.parametrize("flux", ["kepes_tecno_roe", "kepes_tecno_rusanov"])
def test_residual_each_reconstruction(flux, rec):
    m = riemann_mesh(25)
    w = riemann(m.x, m.y)
    prm = T.Params(flux=flux, reconstruct=rec, tvb_M=10.0)
    s, o = _pair(m, prm, w)
    assert rel_l2(s.residual(), o.residual()) < 1e-12


def test_residual_liou_fix_and_evmod():
    m = riemann_mesh(25)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="van_albada", liou_fix=True,
                   ev_alpha=0.3, ev_beta=0.1)
    s, o = _pair(m, prm, riemann(m.x, m.y))
    assert rel_l2(s.residual(), o.residual()) < 1e-12


@pytest.mark.parametrize("path", ["auto", "fused", "graph", "csr"])
def test_liou_fix_runs_on_every_kernel_family(path):
    """The Liou carbuncle fix (src/compute_ssw.cc:7-39, used by KEPES-TeCNO Roe,
    src/kepes_tecno_roe.cc:87) once per step from W^n: 120 steps of the
    Riemann problem, whose shocks set the sensor, on the tiled kernels (the
    sensor rides in the sign of the sound speed's LDS row), the fused stage
    kernel, graph replay and the CSR kernels."""
    m = riemann_mesh(33)
    kw = {"auto": {}, "fused": dict(fused_gradient=True), "graph": dict(graph_steps=8),
          "csr": dict(kernel_path="csr")}[path]
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5, final_time=0.8,
                   liou_fix=True, **kw)
    s, o = _pair(m, prm, riemann(m.x, m.y))
    s.run(120)
    o.run(120)
    assert o.ssw().sum() > 0
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    plain = T.Solver(m, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                                 final_time=0.8, **kw))
    plain.set_state(riemann(m.x, m.y))
    plain.run(120)
    assert not np.array_equal(plain.get_state(), s.get_state())


@pytest.mark.parametrize("bc", ["slip", "farfield", "inlet", "pressure"])
def test_residual_boundary_types(bc):
    st = (1.0, 0.3, 0.1, 1.2)
    m = T.jittered_mesh(21, seed=4, bc={1: (bc, st), 2: ("outlet",), 3: (bc, st), 4: ("outlet",)})
    w = _perturbed_vortex(m) if False else riemann(m.x, m.y)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod")
    s, o = _pair(m, prm, w)
    assert rel_l2(s.residual(), o.residual()) < 1e-12


def test_dt():
    m = riemann_mesh(33)
    s, o = _pair(m, T.Params(cfl=0.5, final_time=0.8), riemann(m.x, m.y))
    a, b = s.compute_dt(), o.compute_dt()
    assert abs(a - b) <= 1e-14 * abs(b)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_vortex_334_steps(path):
    m = vortex_mesh(33)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0,
                   kernel_path=path)
    s, o = _pair(m, prm, vortex(m.x, m.y))
    done, log = s.run(334, want_log=True)
    n, olog = o.run(334, want_log=True)
    assert done[0] == n == 334
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    assert np.allclose(log[:, 1], olog[:, 1], rtol=1e-12, atol=0)   # dt column
    assert np.allclose(log[:, 3], olog[:, 3], rtol=1e-9, atol=0)    # residual norm


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_riemann_334_steps(path):
    m = riemann_mesh(33)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8,
                   kernel_path=path)
    s, o = _pair(m, prm, riemann(m.x, m.y))
    s.run(334)
    o.run(334)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10


def test_rk4_and_fixed_dt():
    m = vortex_mesh(25)
    prm = T.Params(flux="kepec", time_scheme="rk4", cfl=0.0, time_step=0.01, final_time=10.0)
    s, o = _pair(m, prm, vortex(m.x, m.y))
    s.run(50)
    o.run(50)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-11


@pytest.mark.parametrize("path", ["auto", "csr"])
@pytest.mark.parametrize("scheme", ["heuns", "rk4"])
def test_heun_and_rk4_runs_match_oracle(scheme, path):
    """Heun's two-stage scheme and RK4 (tiled and CSR kernels) against the
    oracle, logs included."""
    m = riemann_mesh(33)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8,
                   time_scheme=scheme, kernel_path=path)
    s, o = _pair(m, prm, riemann(m.x, m.y))
    _, lg = s.run(100, want_log=True)
    _, lo = o.run(100, want_log=True)
    assert np.allclose(lg[:, 1:], lo[:, 1:], rtol=1e-12, atol=0)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10


@pytest.mark.parametrize("graph", [0, 3])
def test_rk4_periodic_vortex_tiled_matches_oracle(graph):
    """RK4 with KEPEC on the periodic vortex (the scheme and flux of the
    reference's isentropic-vortex example) on the tiled kernels, eager and
    replayed from HIP graphs."""
    m = vortex_mesh(33)
    prm = T.Params(flux="kepec", reconstruct="first", cfl=0.4, final_time=10.0, time_scheme="rk4",
                   graph_steps=graph)
    s, o = _pair(m, prm, vortex(m.x, m.y))
    for chunk in (7, 13):
        s.run(chunk)
        o.run(chunk)
        assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10


@pytest.mark.parametrize("path,flux,rec", [("auto", "kepec", "first"), ("csr", "kepec", "first"),
                                           ("auto", "kepes_tecno_roe", "minmod")])
def test_navier_stokes_tgv(path, flux, rec):
    """Taylor-Green vortex with viscous and heat fluxes: tiled stage kernels
    (viscous cell fluxes from LDS, viscous time-step term) and the CSR path
    against the oracle, per-step logs included."""
    m = tgv_mesh(25)
    prm = T.Params(model="ns", flux=flux, reconstruct=rec, mu_ref=1.0 / 1600, prandtl=0.72,
                   cfl=0.5, final_time=10.0, kernel_path=path)
    s, o = _pair(m, prm, taylor_green(m.x, m.y))
    # at Mach 0.1 the residual (~1e-3) is a difference of fluxes carrying the
    # background pressure 1/(gamma M^2) ~ 71: the ulp-level differences of the
    # fused multiply-adds in the TeCNO flux grow by that ratio
    assert rel_l2(s.residual(), o.residual()) < (1e-12 if flux == "kepec" else 1e-11)
    _, lg = s.run(100, want_log=True)
    _, lo = o.run(100, want_log=True)
    assert np.allclose(lg[:, 1:3], lo[:, 1:3], rtol=1e-12, atol=0)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    k1, e1 = s.global_integrals()
    k2, e2 = o.global_integrals()
    assert abs(k1 - k2) <= 1e-12 * abs(k2) and abs(e1 - e2) <= 1e-12 * abs(e2)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_final_time_and_error_norms(path):
    m = vortex_mesh(17)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=1.0,
                   time_stamps=2, kernel_path=path)
    s, o = _pair(m, prm, vortex(m.x, m.y))
    done = s.run(100000)
    n = o.run(100000)
    assert done[0] == n
    t, it = s.time()
    assert t == o.time() == 1.0
    ex = vortex(m.x, m.y, 1.0)
    a, b = s.error_norms(ex), o.error_norms(ex)
    assert np.allclose(a, b, rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_deterministic_rerun(path):
    m = riemann_mesh(33)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8,
                   kernel_path=path)
    s = T.Solver(m, prm)
    s.set_state(riemann(m.x, m.y))
    s.run(50)
    a = s.get_state()
    s.set_state(riemann(m.x, m.y))
    s.run(50)
    assert np.array_equal(a, s.get_state())


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_batched_samples_equal_single(path):
    m = vortex_mesh(25)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=0.5,
                   kernel_path=path)
    states = [vortex(m.x, m.y, beta=b) for b in (4.0, 5.0, 6.0)]
    sb = T.Solver(m, prm, n_samples=3)
    for k, w in enumerate(states):
        sb.set_state(w, k)
    sb.run(10000)
    for k, w in enumerate(states):
        s1 = T.Solver(m, prm)
        s1.set_state(w)
        s1.run(10000)
        assert np.array_equal(sb.get_state(k), s1.get_state())


TILED_CASES = [(f, "minmod") for f in FLUXES] + [("kepes_tecno_roe", r) for r in RECS] + \
    [("kepes_tecno_rusanov", r) for r in ("first", "second", "van_albada", "tvb_minmod")]


@pytest.mark.parametrize("flux,rec", TILED_CASES)
def test_tiled_runs_each_scheme(flux, rec):
    """The LDS-tiled stage kernels against the oracle, 20 steps, on a mesh with
    outlet, slip, inlet and farfield sides."""
    st = (1.0, 0.3, 0.1, 1.2)
    m = T.jittered_mesh(23, seed=5, bc={1: ("slip",), 2: ("outlet",), 3: ("inlet", st),
                                        4: ("farfield", st)})
    prm = T.Params(flux=flux, reconstruct=rec, tvb_M=10.0, cfl=0.4, final_time=0.8)
    # smooth state near the boundary values (the sides stay sub-sonic and stable)
    w = np.tile(np.array(st), (m.n_vertex, 1))
    w[:, 0] *= 1 + 0.1 * np.sin(2 * np.pi * m.x) * np.sin(2 * np.pi * m.y)
    w[:, 3] *= 1 + 0.05 * np.cos(2 * np.pi * m.x)
    s, o = _pair(m, prm, w)
    s.run(20)
    o.run(20)
    assert np.all(np.isfinite(o.get_state()))
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-12


@pytest.mark.parametrize("tile", [16, 64, 256])
def test_tiled_tile_sizes_agree_bitwise(tile):
    m = vortex_mesh(41)
    base = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0)
    a = T.Solver(m, base)
    a.set_state(vortex(m.x, m.y))
    a.run(30)
    b = T.Solver(m, T.Params(**{**base.__dict__, "tile_vertices": tile}))
    b.set_state(vortex(m.x, m.y))
    b.run(30)
    assert np.array_equal(a.get_state(), b.get_state())


def test_tiled_then_csr_api_calls_see_the_same_state():
    m = vortex_mesh(25)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0)
    s, o = _pair(m, prm, vortex(m.x, m.y))
    s.run(7)
    o.run(7)
    assert rel_l2(s.residual(), o.residual()) < 1e-11
    assert rel_l2(s.gradients(), o.gradients()) < 1e-11
    s.run(3)
    o.run(3)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-12


def test_monte_carlo_moments_probes_and_integrals():
    """Device reductions over the sample dimension against the oracle's
    restatement of increment_moments on the same sample states (bit-exact:
    same operations in the same order), probe densities against barycentric
    interpolation on the host, and per-sample integrals against the oracle."""
    m = vortex_mesh(33)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0)
    betas = (4.0, 4.5, 5.0, 5.5, 6.0)
    s = T.Solver(m, prm, n_samples=len(betas))
    for k, b in enumerate(betas):
        s.set_state(vortex(m.x, m.y, beta=b), k)
    s.run(12)
    states = np.array([s.get_state(k) for k in range(len(betas))])
    mean, m2, s2 = s.moments()
    om, oq, oz = O.moments(states, np.full(len(betas), 1.4), np.ones(len(betas)))
    assert np.array_equal(mean, om) and np.array_equal(m2, oq) and np.array_equal(s2, oz)
    # probes: density interpolated in the containing triangle
    rng = np.random.default_rng(11)
    px, py = rng.uniform(-4.9, 4.9, 16), rng.uniform(-4.9, 4.9, 16)
    dens = s.probe_density(px, py)
    cells = m.cell_v.reshape(-1, 3)
    rho = states[:, :, 3] / states[:, :, 0]
    for p in range(px.size):
        X, Y = m.x[cells], m.y[cells]
        det = (X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (Y[:, 1] - Y[:, 0])
        l1 = ((px[p] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (py[p] - Y[:, 0])) / det
        l2 = ((X[:, 1] - X[:, 0]) * (py[p] - Y[:, 0]) - (px[p] - X[:, 0]) * (Y[:, 1] - Y[:, 0])) / det
        c = int(np.argmax(np.minimum(np.minimum(1 - l1 - l2, l1), l2)))
        wts = np.array([1 - l1[c] - l2[c], l1[c], l2[c]])
        ref = rho[:, cells[c]] @ wts
        assert np.allclose(dens[:, p], ref, rtol=1e-13)
    with pytest.raises(T.TsmError):
        s.probe_density(np.array([50.0]), np.array([0.0]))
    ke, ent = s.global_all()
    o = O.Oracle(m, prm)
    for k in range(len(betas)):
        o.set_state(states[k])
        ko, eo = o.global_integrals()
        assert abs(ke[k] - ko) <= 1e-12 * abs(ko) and abs(ent[k] - eo) <= 1e-12 * abs(eo)


@pytest.mark.parametrize("path,chunk", [("auto", 100000), ("csr", 100000), ("auto", 1), ("csr", 1)])
def test_dt_log_matches_oracle(path, chunk):
    """Per-step (dt, t, res) logs against the oracle at 1e-12, whole runs in
    one call and one step per call (as the file driver advances)."""
    m = vortex_mesh(17)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=1.0,
                   time_stamps=2, kernel_path=path)
    s, o = _pair(m, prm, vortex(m.x, m.y, box=None))
    n, olog = o.run(100000, want_log=True)
    rows = []
    while True:
        done, lg = s.run(chunk, want_log=True)
        if done[0] == 0:
            break
        rows.append(lg[:done[0]])
    glog = np.vstack(rows)
    assert glog.shape == olog.shape
    d = np.abs(glog[:, 1:] - olog[:, 1:]) / np.abs(olog[:, 1:])
    print("max rel diff per column (dt, t, res):", d.max(axis=0), "first bad row:",
          int(np.argmax(d.max(axis=1) > 1e-12)))
    assert np.all(d <= 1e-12)


@pytest.mark.parametrize("flux,rec", [("kepes_tecno_rusanov", "minmod"), ("kepes_tecno_roe", "minmod"),
                                      ("kepes_tecno_roe", "van_albada"), ("kepes_tecno_rusanov", "eno2"),
                                      ("kepes_tecno_roe", "second")])
def test_fused_two_ring_stage_equals_separate_gradient_pass(flux, rec):
    """On a mesh without periodic sets the stage kernel forms the gradients
    itself (two halo rings); the result must equal the separate gradient pass
    bit for bit, boundary corrections of every side type included."""
    st = (1.0, 0.3, 0.1, 1.2)
    m = T.jittered_mesh(29, seed=6, bc={1: ("slip",), 2: ("outlet",), 3: ("inlet", st),
                                        4: ("farfield", st)})
    w = np.tile(np.array(st), (m.n_vertex, 1))
    w[:, 0] *= 1 + 0.1 * np.sin(2 * np.pi * m.x) * np.sin(2 * np.pi * m.y)
    w[:, 3] *= 1 + 0.05 * np.cos(2 * np.pi * m.x)
    out = []
    for fused in (True, False):
        prm = T.Params(flux=flux, reconstruct=rec, cfl=0.4, final_time=0.8, fused_gradient=fused)
        s = T.Solver(m, prm)
        s.set_state(w)
        s.run(15)
        out.append(s.get_state())
    assert np.array_equal(out[0], out[1])


def test_autotuned_shape_gives_identical_results():
    from tensum_amd import autotune as AT
    m = riemann_mesh(65)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8)
    w = riemann(m.x, m.y)
    shapes = AT.candidates(tiles=(64, 128, 256), fu_blocks=(0, 512))
    best, table = AT.autotune(m, prm, w, steps=3, shapes=shapes)
    assert len(table) >= 3 and all(us > 0 for _, us in table)
    out = []
    for p in (prm, best, T.Params(**{**prm.__dict__, "tile_vertices": 64, "block_fu": 512})):
        s = T.Solver(m, p)
        s.set_state(w)
        s.run(12)
        out.append(s.get_state())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


@pytest.mark.parametrize("mesh_kind", ["riemann", "vortex"])
def test_graph_replays_equal_eager_steps(mesh_kind):
    """Steps replayed from a captured HIP graph (5 steps per graph) equal
    host-launched steps bit for bit, replay after replay, with eager work
    (state reads, residual calls, a second sample's reset) between replays."""
    if mesh_kind == "riemann":
        m = riemann_mesh(41)
        prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8)
        w = riemann(m.x, m.y)
    else:
        m = vortex_mesh(33)
        prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0)
        w = vortex(m.x, m.y)
    g = T.Solver(m, T.Params(**{**prm.__dict__, "graph_steps": 5}), n_samples=2)
    e = T.Solver(m, prm, n_samples=2)
    for s in (g, e):
        s.set_state(w, 0)
        s.set_state(w, 1)
    for chunk in (12, 5, 10, 3):
        dg, de = g.run(chunk), e.run(chunk)
        assert np.array_equal(dg, de)
        for k in (0, 1):
            assert np.array_equal(g.get_state(k), e.get_state(k))
            assert g.time(k) == e.time(k)
        g.residual(0)  # eager CSR work between replays
        e.residual(0)
    w2 = g.get_state(0)
    g.set_state(w2, 1)
    e.set_state(w2, 1)
    g.run(10)
    e.run(10)
    assert np.array_equal(g.get_state(1), e.get_state(1))


def _fan_mesh(n=12):
    """Half disk around the centre of its diameter: an inner fan of n
    triangles (so the centre, a boundary vertex, has n cells, more than one
    lane group of the boundary-gradient kernel takes at once) and an outer band
    of 2n triangles."""
    th = np.pi * np.arange(n + 1) / n
    x = np.concatenate([[0.0], 0.5 * np.cos(th), np.cos(th)])
    y = np.concatenate([[0.0], 0.5 * np.sin(th), np.sin(th)])
    inn = 1 + np.arange(n + 1)
    out = n + 2 + np.arange(n + 1)
    cells = [(0, inn[k], inn[k + 1]) for k in range(n)]
    for k in range(n):
        cells += [(inn[k], out[k], out[k + 1]), (inn[k], out[k + 1], inn[k + 1])]
    # boundary counter-clockwise: the diameter from (-1,0) to (1,0), then the arc
    bf = [(out[n], inn[n], 1), (inn[n], 0, 1), (0, inn[0], 1), (inn[0], out[0], 1)]
    bf += [(out[k], out[k + 1], 2) for k in range(n)]
    return T.Mesh(x=x, y=y, bface_tag=np.array([t for _, _, t in bf], np.int32),
                  bface_v=np.array([[a, b] for a, b, _ in bf], np.int32).ravel(),
                  cell_v=np.array(cells, np.int32).ravel(),
                  periodic_ptr=np.zeros(0, np.int32), periodic_v=np.zeros(0, np.int32),
                  bc={1: ("slip",), 2: ("outlet",)})


def test_boundary_vertex_with_many_cells():
    m = _fan_mesh(12)
    st = np.array([1.0, 0.2, 0.1, 1.0])
    w = np.tile(st, (m.n_vertex, 1))
    w[:, 3] *= 1 + 0.05 * np.cos(3 * m.x) * (1 + m.y)
    for fused in (False, True):
        prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.3, final_time=1.0,
                       fused_gradient=fused)
        s, o = _pair(m, prm, w)
        assert rel_l2(s.gradients(), o.gradients()) < 1e-12
        s.run(6)
        o.run(6)
        assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-12


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_step_matches_run_and_lands_on_target(path):
    """tsm_step advances exactly like tsm_run; with a target time it clamps the
    step so the sample's time lands on it."""
    m = vortex_mesh(17)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0,
                   kernel_path=path)
    a, b = T.Solver(m, prm), T.Solver(m, prm)
    for s in (a, b):
        s.set_state(vortex(m.x, m.y))
    _, log = a.run(5, want_log=True)
    for k in range(5):
        dt, res = b.step()
        assert dt[0] == log[k, 1] and res[0] == log[k, 3]
    assert np.array_equal(a.get_state(), b.get_state())
    t0 = b.time()[0]
    target = t0 + 0.3 * log[-1, 1]
    b.step(target)
    assert abs(b.time()[0] - target) <= 2e-16 * target  # t + (target - t)
    b.step()
    assert b.time()[0] > target
