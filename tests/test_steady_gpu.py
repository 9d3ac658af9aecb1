"""Steady mode (local time stepping, src/fv.cc:1499-1621, :2230-2395) on the GPU
against the oracle: each vertex advances with its own dt_i = cfl A_i /
lambda_i, t advances by min dt_i, and the loop stops after max_iter steps or
once the normalised residual is <= min_residue.  Both kernel families (the
tiled stage kernels write the next step's dt_i in the last stage; the CSR
kernels in the time-step pass), periodic and wall-bounded meshes, Euler and
Navier-Stokes.  Tolerances as for unsteady runs: logs to 1e-12 relative,
states to rel-L2 1e-10 in conserved variables.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2, taylor_green, tgv_mesh, vortex, vortex_mesh

pytestmark = pytest.mark.gpu

FREE = (1.0, 0.5, 0.0, 1.0)


def _channel(n, seed=3):
    return T.jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=seed,
                           bc={1: ("slip",), 3: ("slip",), 2: ("outlet",), 4: ("farfield", FREE)})


def _bump(x, y):
    g = np.exp(-40.0 * ((x - 0.45) ** 2 + (y - 0.5) ** 2))
    return np.stack([1.0 + 0.1 * g, 0.5 + 0.05 * g, 0.03 * g, 1.0 + 0.2 * g], 1)


def _steady(**kw):
    base = dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5, time_mode="steady",
                max_iter=10 ** 6)
    return T.Params(**{**base, **kw})


def _compare(s, o, nsteps, res_tol=1e-12):
    done, lg = s.run(nsteps, want_log=True)
    n, lo = o.run(nsteps, want_log=True)
    assert done[0] == n
    lg = lg[:n]
    assert np.array_equal(lg[:, 0], lo[:, 0])
    d = np.abs(lg[:, 1:3] - lo[:, 1:3]) / np.abs(lo[:, 1:3])
    assert d.max() <= 1e-12, d.max()
    r = np.abs(lg[:, 3] - lo[:, 3]) / np.abs(lo[:, 3])
    assert r.max() <= res_tol, r.max()
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    return lg


@pytest.mark.parametrize("path,flux,rec", [("auto", "kepes_tecno_roe", "minmod"),
                                           ("csr", "kepes_tecno_roe", "minmod"),
                                           ("auto", "kepes_tecno_rusanov", "van_albada"),
                                           ("auto", "kepec", "first")])
def test_steady_channel_matches_oracle(path, flux, rec):
    m = _channel(33)
    prm = _steady(flux=flux, reconstruct=rec, kernel_path=path)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    w = _bump(m.x, m.y)
    s.set_state(w)
    o.set_state(w)
    _compare(s, o, 60)


def test_steady_fused_gradient_stage():
    m = _channel(33)
    prm = _steady(fused_gradient=True)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(_bump(m.x, m.y))
    o.set_state(_bump(m.x, m.y))
    _compare(s, o, 40)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_steady_periodic_vortex_matches_oracle(path):
    m = vortex_mesh(33)
    prm = _steady(cfl=0.8, kernel_path=path)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(vortex(m.x, m.y))
    o.set_state(vortex(m.x, m.y))
    _compare(s, o, 50)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_steady_navier_stokes_matches_oracle(path):
    m = tgv_mesh(25)
    prm = _steady(model="ns", flux="kepec", reconstruct="first", mu_ref=1.0 / 400,
                  kernel_path=path)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(taylor_green(m.x, m.y))
    o.set_state(taylor_green(m.x, m.y))
    _compare(s, o, 40)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_steady_stops_like_the_oracle(path):
    """max_iter and min_residue end the loop at the oracle's step, in chunks."""
    m = _channel(17)
    for kw in (dict(max_iter=23), dict(min_residue=1e-3, cfl=0.8)):
        prm = _steady(kernel_path=path, **kw)
        s, o = T.Solver(m, prm), O.Oracle(m, prm)
        s.set_state(_bump(m.x, m.y))
        o.set_state(_bump(m.x, m.y))
        n = o.run(100000)
        total = 0
        while True:
            d = int(s.run(7)[0])
            if d == 0:
                break
            total += d
        assert total == n and s.status()["done"]
        assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10


def test_steady_residual_norms_and_t_target():
    m = _channel(17)
    s = T.Solver(m, _steady())
    s.set_state(_bump(m.x, m.y))
    s.run(1)
    r1 = s.residual_norms()
    _, lg = s.run(5, want_log=True)
    r6 = s.residual_norms()
    # res_total is the norm of the component norms over the first step's
    assert np.isclose(np.linalg.norm(r6) / np.linalg.norm(r1), lg[-1, 3], rtol=1e-14)
    assert np.all(r1 > 0)
    with pytest.raises(T.TsmError, match="t_target"):
        s.step(1.0)
    dt, res = s.step()
    assert dt[0] > 0 and res[0] > 0


def test_steady_batched_samples_and_graph_replay():
    """Per-sample local steps: a batch of three equals three single runs, and
    graph replay equals host-launched steps bit for bit."""
    m = _channel(25)
    states = [_bump(m.x, m.y) * np.array([1.0, a, 1.0, 1.0]) for a in (0.9, 1.0, 1.1)]
    out = {}
    for graph in (0, 4):
        sb = T.Solver(m, _steady(graph_steps=graph), n_samples=3)
        for k, w in enumerate(states):
            sb.set_state(w, k)
        sb.run(20)
        out[graph] = [sb.get_state(k) for k in range(3)]
    for k, w in enumerate(states):
        s1 = T.Solver(m, _steady())
        s1.set_state(w)
        s1.run(20)
        assert np.array_equal(out[0][k], s1.get_state())
        assert np.array_equal(out[4][k], s1.get_state())
