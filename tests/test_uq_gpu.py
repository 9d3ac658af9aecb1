"""Monte Carlo layer on the GPU: batched samples with per-sample material,
levels advanced together, multilevel estimates and the surrogate control
variate on solver-generated quantities of interest."""
import numpy as np
import pytest

import tensum_amd as T
from tensum_amd import uq
from tensum_amd.surrogate import Surrogate, control_variate, _alpha

pytestmark = pytest.mark.gpu


def test_run_group_equals_separate_runs():
    p = uq.vortex_problem(final_time=0.3)
    inp = uq.sample_inputs(3, p.bounds, seed=2)
    a1 = uq.make_batch(p, p.mesh(17), inp)
    a2 = uq.make_batch(p, p.mesh(25), inp)
    uq.advance([a1, a2], chunk=16)
    for n, a in ((17, a1), (25, a2)):
        b = uq.make_batch(p, p.mesh(n), inp)
        b.run(100000)
        for k in range(3):
            assert np.array_equal(a.get_state(k), b.get_state(k))
            assert abs(a.time(k)[0] - 0.3) <= 1e-15


def test_reused_context_gives_the_results_of_fresh_ones():
    """plain_mc reuses one context for every batch of its size (set_state
    restarts each sample's clock); the QoIs equal those of a fresh context per
    batch bit for bit, including a smaller last batch."""
    p = uq.vortex_problem(final_time=0.4)
    u, Q = uq.plain_mc(p, 17, 10, seed=3, batch=4)
    m = p.mesh(17)
    for b0 in range(0, 10, 4):
        part = {k: v[b0:b0 + 4] for k, v in u.items()}
        s = uq.make_batch(p, m, part)
        uq.advance([s])
        assert np.array_equal(uq.qoi(p, s), Q[b0:b0 + 4])
        s.close()


def test_reference_in_chunks_is_the_reference_in_one_run():
    """reference_chunk: consecutive chunks of the plain-MC reference's Sobol
    sequence give the QoIs of the whole reference solved in one run, bit for
    bit, and cv_experiment takes them in place of its own reference."""
    p = uq.vortex_problem(final_time=0.4)
    Q_all, _ = uq.reference_chunk(p, 17, 0, 10, seed=2, fine_batch=4)
    _, Q_one = uq.plain_mc(p, 17, 10, seed=2 + 1000, batch=4)
    assert np.array_equal(Q_all, Q_one)
    parts = [uq.reference_chunk(p, 17, a, n, seed=2, fine_batch=4)[0] for a, n in ((0, 3), (3, 7))]
    assert np.array_equal(np.vstack(parts), Q_all)
    with pytest.raises(ValueError):  # a reference of the wrong size is refused
        uq.cv_experiment(p, 17, 17, 8, 0, 4, 64, 12, epochs=1, seed=2, ref=(Q_all, 1.0))


def test_batched_samples_match_one_at_a_time_with_material():
    p = uq.vortex_problem(final_time=0.2)
    inp = uq.sample_inputs(4, p.bounds, seed=3)
    m = p.mesh(17)
    s = uq.make_batch(p, m, inp)
    uq.advance([s])
    Q = uq.qoi(p, s)
    for k in range(4):
        one = uq.make_batch(p, m, {n: v[k:k + 1] for n, v in inp.items()})
        uq.advance([one])
        assert np.array_equal(uq.qoi(p, one)[0], Q[k])
        assert np.array_equal(one.get_state(0), s.get_state(k))


def test_mlmc_agrees_with_plain_mc_on_the_finest_level():
    p = uq.vortex_problem(final_time=0.5)
    est = uq.mlmc(p, levels=(17, 33), samples=(64, 16), seed=4)
    _, Q = uq.plain_mc(p, 33, 64, seed=9)
    mq, vq = uq.moments(Q)
    se = np.sqrt(vq / 64 + est.mean_se ** 2)
    z = np.abs(est.mean - mq) / np.maximum(se, 1e-300)
    assert np.mean(z < 4) > 0.95
    assert np.all(np.isfinite(est.var))


def test_surrogate_control_variate_beats_plain_mc_on_solver_qoi():
    """Isentropic vortex with uncertain strength, centre, gamma and Mach on a
    33^2 mesh (t = 0.5): the surrogate is trained on 512 solver samples, its
    control coefficients fitted on 32 independent ones, and it then serves as
    control variate for 48 further solver samples plus 8192 surrogate
    evaluations.  The QoIs depend smoothly on the inputs, so the variance of
    the estimate must drop at least 4x against the plain mean of the same 48
    solver samples (integrated kinetic energy and entropy, and the median over
    the probes that vary)."""
    p = uq.vortex_problem(final_time=0.5)

    def unit(inp):
        return np.stack([inp[b[0]] for b in p.bounds], 1)

    u_tr, Q_tr = uq.plain_mc(p, 33, 512, seed=11)
    s = Surrogate(5, p.n_qoi, max_batch=512, seed=2)
    s.fit(unit(u_tr), Q_tr, epochs=300, batch=64, lr=2e-3)
    u_a, Q_a = uq.plain_mc(p, 33, 32, seed=14)
    G_a = s(unit(u_a))
    u_p, Q_p = uq.plain_mc(p, 33, 48, seed=12)
    U_m = unit(uq.sample_inputs(8192, p.bounds, seed=13))
    est = control_variate(Q_p, s(unit(u_p)), s(U_m), alpha=_alpha(Q_a, G_a),
                          alpha2=_alpha(Q_a ** 2, G_a ** 2))
    varying = est.plain_mean_se[:-2] > 1e-6 * np.abs(est.mean[:-2])
    vr_probes = np.median(est.variance_reduction[:-2][varying])
    print("variance reduction (probes median, KE, entropy):", vr_probes, est.variance_reduction[-2:])
    assert np.all(est.variance_reduction[-2:] > 4.0) and vr_probes > 4.0


def test_exploded_samples_are_refused():
    p = uq.vortex_problem(final_time=0.05)
    inp = uq.sample_inputs(2, p.bounds, seed=1)
    s = uq.make_batch(p, p.mesh(17), inp)
    w = s.get_state(1)
    w[:, 3] *= 1e9  # a state whose density leaves any physical range
    s.set_state(w, 1)
    with pytest.raises(FloatingPointError):
        uq.qoi(p, s)


def test_riemann_control_variate_trained_on_a_coarse_mesh():
    """The four-quadrant Riemann problem with uncertain split (U(0.75,0.85)^2)
    and quadrant pressures (x U(0.9,1.1)), first-order KEPES-TeCNO Rusanov
    (the second-order scheme takes part of these samples to non-physical
    states), t = 0.02, at reduced size: the
    surrogate is trained on 256 solves on the 65^2 mesh, its coefficients are
    fitted on those training samples (independent of the paired ones, so the
    estimate stays unbiased), and it serves as control variate for 48 paired
    solves on the 129^2 mesh plus 8192 surrogate evaluations.  The integrated
    kinetic energy and entropy depend smoothly on the inputs: their variance
    must drop at least 4x against the plain mean of the 48 fine solves, and
    the estimate must agree with a 256-sample plain-MC reference at 129^2
    within its standard error."""
    p = uq.riemann_problem(final_time=0.02, reconstruct="first")

    def unit(inp):
        return np.stack([inp[b[0]] for b in p.bounds], 1)

    u_tr, Q_tr = uq.plain_mc(p, 65, 256, seed=21)
    s = Surrogate(6, p.n_qoi, max_batch=512, seed=3)
    s.fit(unit(u_tr), Q_tr, epochs=300, batch=64, lr=2e-3)
    G_tr = s(unit(u_tr))
    u_p, Q_p = uq.plain_mc(p, 129, 48, seed=22)
    U_m = unit(uq.sample_inputs(8192, p.bounds, seed=23))
    est = control_variate(Q_p, s(unit(u_p)), s(U_m), alpha=_alpha(Q_tr, G_tr),
                          alpha2=_alpha(Q_tr ** 2, G_tr ** 2))
    print("variance reduction (KE, entropy):", est.variance_reduction[-2:], flush=True)
    assert np.all(est.variance_reduction[-2:] > 4.0)
    _, Q_ref = uq.plain_mc(p, 129, 256, seed=24)
    m_ref, v_ref = uq.moments(Q_ref)
    se = np.sqrt(est.mean_se[-2:] ** 2 + v_ref[-2:] / 256)
    assert np.all(np.abs(est.mean[-2:] - m_ref[-2:]) < 4 * se)
