"""The gradients the reference's outputs read after a step (those of the last
stage's input state, src/fv.cc:1476, :2295-2304) and the skin-friction and
heat-flux surfaces (src/output_skin_friction.cc, src/output_heatflux.cc) on
the GPU against the oracle: tiled, fused-gradient and CSR kernels, the three
Runge-Kutta schemes, Euler and Navier-Stokes, batched samples that stop at
different steps, and step graphs captured before the state was kept.
Tolerances as for runs: rel-L2 1e-10.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import rel_l2

pytestmark = pytest.mark.gpu

ST = (1.0, 0.4, 0.05, 1.0)
EULER_BC = {1: ("slip",), 3: ("slip",), 4: ("inlet", ST), 2: ("outlet",)}
NS_BC = {1: ("noslip", (1.2, 0.0, 0.0, 1.0)), 3: ("noslip", (1.0, 0.1, 0.0, 1.0), True),
         4: ("farfield", ST), 2: ("outlet",)}


def _state(m, amp=0.1):
    return np.stack([1.0 + amp * np.sin(np.pi * m.x) * np.cos(np.pi * m.y), 0.4 + amp * m.y,
                     0.05 * np.cos(np.pi * m.x), 1.0 + 0.5 * amp * np.cos(np.pi * (m.x + m.y))], 1)


def _prm(**kw):
    return T.Params(**{**dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                              final_time=10.0), **kw})


CASES = {
    "euler_tiled_ssprk3": (EULER_BC, _prm()),
    "euler_fused_ssprk3": (EULER_BC, _prm(fused_gradient=True)),
    "euler_tiled_rk4": (EULER_BC, _prm(time_scheme="rk4")),
    "euler_tiled_heun": (EULER_BC, _prm(time_scheme="heuns")),
    "euler_csr_ssprk3": (EULER_BC, _prm(kernel_path="csr")),
    "euler_csr_rk4": (EULER_BC, _prm(kernel_path="csr", time_scheme="rk4")),
    "ns_tiled_ssprk3": (NS_BC, _prm(model="ns", mu_ref=0.01)),
    "ns_csr_heun": (NS_BC, _prm(model="ns", mu_ref=0.01, kernel_path="csr", time_scheme="heuns")),
    "ns_steady": (NS_BC, _prm(model="ns", mu_ref=0.01, time_mode="steady", max_iter=10 ** 6)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_stage_gradients_and_surfaces_match_oracle(case):
    bc, prm = CASES[case]
    m = T.jittered_mesh(25, seed=5, bc=bc)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.keep_stage_state()
    s.set_state(_state(m))
    o.set_state(_state(m))
    assert np.array_equal(s.stage_gradients(), s.gradients())   # before a step
    tags = [1, 3]
    for steps in (7, 13):
        s.run(steps)
        o.run(steps)
        g = o.stage_gradients()
        assert rel_l2(s.stage_gradients(), g) <= 1e-10
        assert rel_l2(s.gradients(), g) > 1e-6                   # not the current state's
        if prm.model == "ns":
            a, b = s.surface(tags), o.surface(tags)
            assert a.shape == b.shape == (int(np.isin(m.bface_tag, tags).sum()), 2)
            assert rel_l2(a, b) <= 1e-10
            # a Navier-Stokes force evaluation recomputes the oracle's dE on the
            # current state: the current-state surface then matches
            o.forces(tags, force_index=0)
            assert rel_l2(s.surface(tags, stage=False), o.surface(tags)) <= 1e-10


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_batched_samples_keep_their_own_last_step(path):
    """Samples with different states take different time steps and stop at
    the final time after different step counts; each keeps the last-stage
    state of its own last step."""
    m = T.jittered_mesh(21, seed=6, bc=EULER_BC)
    prm = _prm(final_time=0.05, kernel_path=path)
    amps = (0.05, 0.3, 0.6)
    orcs = []
    for a in amps:
        o = O.Oracle(m, prm)
        o.set_state(_state(m, a))
        orcs.append((o, o.run(100000)))
    assert len(set(n for _, n in orcs)) > 1
    s = T.Solver(m, prm, n_samples=len(amps))
    s.keep_stage_state()
    for k, a in enumerate(amps):
        s.set_state(_state(m, a), k)
    done = s.run(max(n for _, n in orcs) + 3)
    for k, (o, n) in enumerate(orcs):
        assert done[k] == n
        assert rel_l2(s.stage_gradients(k), o.stage_gradients()) <= 1e-10
        assert rel_l2(s.get_state(k), o.get_state()) <= 1e-10


def test_graph_captured_before_keeping_is_recaptured():
    m = T.jittered_mesh(25, seed=7, bc=EULER_BC)
    prm = _prm(graph_steps=4)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(_state(m))
    o.set_state(_state(m))
    s.run(8)
    s.keep_stage_state()
    s.run(8)
    o.run(16)
    assert rel_l2(s.stage_gradients(), o.stage_gradients()) <= 1e-10
    s.keep_stage_state(False)
    with pytest.raises(T.TsmError):
        s.stage_gradients()


def test_stage_gradients_need_keeping():
    m = T.jittered_mesh(9, seed=7, bc=EULER_BC)
    s = T.Solver(m, _prm())
    s.set_state(_state(m))
    with pytest.raises(T.TsmError):
        s.stage_gradients()
    with pytest.raises(T.TsmError):
        s.surface([1])
    assert s.surface([1], stage=False).shape == ((np.asarray(m.bface_tag) == 1).sum(), 2)
