"""Monte Carlo and multilevel Monte Carlo runs on the batched solver.

The reference runs one sample after another (src/fv.cc:2440-2459), reads its
random inputs from a file (src/fv.cc:502-551) and accumulates per-vertex
moments with Welford updates (src/fv.cc:595-694) that a final pass merges
across ranks (src/reduce_moments.cc:14-220).  Here every sample of a level is
one slot of a single libtsm context (sample index = grid dimension of every
kernel, per-sample dt and stop time), the levels of a multilevel run advance
together on their own streams, inputs come from the scrambled Sobol sampler
and the quantities of interest (probe densities, integrated entropy) are
reduced on the device.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import Params, Solver, TsmError, jittered_mesh, run_group, run_packed, sobol

# ---------------------------------------------------------------------------
# uncertain inputs
# ---------------------------------------------------------------------------
VORTEX_BOUNDS = (("beta", 4.0, 6.0), ("xc", -0.5, 0.5), ("yc", -0.5, 0.5),
                 ("gamma", 1.3, 1.5), ("mach", 0.4, 0.6))
RIEMANN_BOUNDS = (("sx", 0.75, 0.85), ("sy", 0.75, 0.85), ("p_ne", 0.9, 1.1),
                  ("p_nw", 0.9, 1.1), ("p_sw", 0.9, 1.1), ("p_se", 0.9, 1.1))


def map_inputs(u, bounds):
    """Unit-cube points (n, d) -> dict of uniformly distributed inputs."""
    u = np.asarray(u, dtype=np.float64)
    return {name: lo + (hi - lo) * u[:, k] for k, (name, lo, hi) in enumerate(bounds)}


def sample_inputs(n, bounds, seed=0, skip=0):
    return map_inputs(sobol(n, len(bounds), seed=seed, skip=skip), bounds)


# ---------------------------------------------------------------------------
# initial states
# ---------------------------------------------------------------------------
def vortex_state(x, y, beta, xc, yc, gamma, mach, t=0.0, box=10.0):
    """(T,u,v,p) of the isentropic vortex on the periodic box [-box/2, box/2]^2:
    freestream rho = p = T = 1 (R = 1), velocity of Mach `mach` along (1,1),
    strength beta, centre (xc, yc) at t = 0; the same expressions as the
    reference's examples/Euler/isentropic_vortex/param.in:72-96 with a shifted
    centre and a scaled freestream velocity."""
    speed = mach * np.sqrt(gamma)
    ux = uy = speed / np.sqrt(2.0)
    xr = ((x - xc - ux * t) + 0.5 * box) % box - 0.5 * box
    yr = ((y - yc - uy * t) + 0.5 * box) % box - 0.5 * box
    r2 = xr * xr + yr * yr
    Tm = 1.0 - (gamma - 1.0) * (beta * beta / (8.0 * gamma * np.pi ** 2)) * np.exp(1.0 - r2)
    e = beta / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    return np.stack([Tm, ux - e * yr, uy + e * xr, Tm ** (gamma / (gamma - 1.0))], 1)


RIEMANN_QUADRANTS = {  # (rho, u, v, p) NE, NW, SW, SE
    "ne": (1.5, 0.0, 0.0, 1.5), "nw": (0.5323, 1.206, 0.0, 0.3),
    "sw": (0.138, 1.206, 1.206, 0.029), "se": (0.5323, 0.0, 1.206, 0.3)}


def riemann_state(x, y, sx, sy, p_ne, p_nw, p_sw, p_se):
    """(T,u,v,p) of the four-quadrant Riemann problem with the split at (sx, sy)
    and each quadrant's pressure scaled (R = 1)."""
    east, north = x >= sx, y >= sy
    out = np.empty((x.size, 4))
    for key, mask, sc in (("ne", east & north, p_ne), ("nw", ~east & north, p_nw),
                          ("sw", ~east & ~north, p_sw), ("se", east & ~north, p_se)):
        rho, u, v, p = RIEMANN_QUADRANTS[key]
        out[mask] = (p * sc / rho, u, v, p * sc)
    return out


def probe_grid(lo, hi, n=8):
    """n x n probe points at the centres of an n x n grid of [lo, hi]^2."""
    c = lo + (np.arange(n) + 0.5) * (hi - lo) / n
    px, py = np.meshgrid(c, c, indexing="xy")
    return px.ravel(), py.ravel()


# ---------------------------------------------------------------------------
# problems
# ---------------------------------------------------------------------------
@dataclass
class Problem:
    """A parametrised workload: mesh builder, solver settings, inputs -> states."""
    name: str
    bounds: tuple
    params: Params
    mesh: callable
    state: callable
    gamma: callable
    probes: tuple
    n_qoi: int = field(default=0)

    def __post_init__(self):
        self.n_qoi = self.probes[0].size + 2


def vortex_problem(final_time=10.0):
    prm = Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=final_time)
    return Problem(
        "vortex", VORTEX_BOUNDS, prm,
        lambda n: jittered_mesh(n, box=(-5.0, 5.0, -5.0, 5.0), jitter=0.2, seed=0, periodic=True),
        lambda m, p: vortex_state(m.x, m.y, p["beta"], p["xc"], p["yc"], p["gamma"], p["mach"]),
        lambda p: p["gamma"], probe_grid(-4.375, 4.375))


def riemann_problem(final_time=0.2, reconstruct="minmod"):
    """The four-quadrant Riemann problem with uncertain split and quadrant
    pressures.  With independently scaled pressures the quadrant interfaces
    no longer carry single elementary waves and some samples drive the
    second-order (minmod) scheme to non-physical states early (measured: 4 %
    of the samples on 257^2 by t = 0.01, 37 % on 1025^2 by t = 0.005, as the
    reference's scheme, which is not positivity preserving, does on the
    unperturbed problem later); first-order reconstruction keeps every sample
    physical to t = 0.2 (257^2) and t = 0.05 (1025^2)."""
    prm = Params(flux="kepes_tecno_rusanov", reconstruct=reconstruct, cfl=0.5,
                 final_time=final_time)
    return Problem(
        "riemann", RIEMANN_BOUNDS, prm,
        lambda n: jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=1, periodic=False,
                                bc={t: ("outlet",) for t in (1, 2, 3, 4)}),
        lambda m, p: riemann_state(m.x, m.y, p["sx"], p["sy"], p["p_ne"], p["p_nw"], p["p_sw"],
                                   p["p_se"]),
        lambda p: np.full(np.size(p["sx"]), 1.4), probe_grid(0.05, 0.95))


# ---------------------------------------------------------------------------
# batched solves
# ---------------------------------------------------------------------------
def _one(inputs, k):
    return {name: float(v[k]) for name, v in inputs.items()}


def make_batch(problem: Problem, mesh, inputs, solver=None):
    """One Solver holding every sample of `inputs` (dict of arrays): `solver`
    when given (a context of the same mesh and sample count, whose samples
    then start again from t = 0: set_state resets each sample's clock), else a
    new one."""
    n = len(next(iter(inputs.values())))
    s = solver if solver is not None else Solver(mesh, problem.params, n_samples=n)
    assert s.n_samples == n
    gam = np.broadcast_to(problem.gamma(inputs), (n,))
    for k in range(n):
        s.set_material(k, float(gam[k]), 1.0)
        s.set_state(problem.state(mesh, _one(inputs, k)), k)
    return s


def qoi(problem: Problem, solver: Solver, max_density=1e6):
    """(n_samples, n_probe + 2): probe densities, kinetic energy, entropy
    integral.  Raises FloatingPointError naming the samples whose state is not
    finite or whose density left [0, max_density] (a run that blew up)."""
    rho = solver.probe_density(*problem.probes)
    ke, ent = solver.global_all()
    Q = np.hstack([rho, ke[:, None], ent[:, None]])
    bad = ~np.all(np.isfinite(Q), axis=1) | np.any((rho < 0) | (rho > max_density), axis=1)
    if np.any(bad):
        raise FloatingPointError(f"{problem.name}: samples {np.flatnonzero(bad).tolist()} "
                                 "have a non-finite or non-physical state")
    return Q


def advance(solvers, chunk=64, max_chunks=100000, packed=False, log=None, log_every=30.0):
    """Runs every solver until all its samples reached final_time: the
    solvers' own launches interleaved on their streams, or (packed=True) one
    launch per kernel spanning every solver where their scheme allows it
    (Euler, SSP-RK3, tiled kernels, one flux and reconstruction; the same
    results bit for bit).  Measured on the four-level vortex batch (65^2 to
    513^2 with 4096/1024/256/64 samples) the streams are faster: 6.2e9 against
    5.1e9 vertex-stage updates/s, as their gradient and flux passes of
    different levels overlap.  log: called with a progress line every
    log_every seconds."""
    step = run_packed if packed else run_group
    t_last = time.perf_counter()
    for k in range(max_chunks):
        if log and time.perf_counter() - t_last > log_every:
            t_last = time.perf_counter()
            log(f"advanced {k * chunk} steps")
        try:
            done = step(solvers, chunk)
        except TsmError as e:
            if step is run_group or "packed" not in str(e) and "share" not in str(e):
                raise
            step = run_group
            continue
        if not np.any(done):
            return
    raise RuntimeError("samples did not reach final_time")


def plain_mc(problem: Problem, n_mesh, n_samples, seed=0, skip=0, batch=None, log=None):
    """Plain Monte Carlo on one mesh: QoI per sample, in sample order, solved
    `batch` samples at a time (all at once by default)."""
    mesh = problem.mesh(n_mesh)
    inputs = sample_inputs(n_samples, problem.bounds, seed=seed, skip=skip)
    batch = batch or n_samples
    out = []
    s = None  # one context serves every batch of its size (set-up is paid once)
    for b0 in range(0, n_samples, batch):
        part = {k: v[b0:b0 + batch] for k, v in inputs.items()}
        nb = len(next(iter(part.values())))
        if s is not None and s.n_samples != nb:
            s.close()
            s = None
        s = make_batch(problem, mesh, part, s)
        advance([s], log=log)
        out.append(qoi(problem, s))
        if log:
            log(f"{min(b0 + batch, n_samples)}/{n_samples} samples at {n_mesh}^2")
    if s is not None:
        s.close()
    return inputs, np.vstack(out)


@dataclass
class Estimate:
    mean: np.ndarray
    var: np.ndarray
    mean_se: np.ndarray        # standard error of the mean estimate
    seconds: float
    detail: dict


def moments(Q):
    """Sample mean and unbiased variance per column (Welford order, as the
    reference's increment_moments + reduce_moments)."""
    Q = np.asarray(Q, dtype=np.float64)
    M = Q[0].copy()
    M2 = np.zeros_like(M)
    for k in range(1, Q.shape[0]):
        d = Q[k] - M
        M += d / (k + 1)
        M2 += d * (Q[k] - M)
    return M, M2 / max(Q.shape[0] - 1, 1)


def mlmc(problem: Problem, levels, samples, seed=0, log=None):
    """Multilevel Monte Carlo: E[Q_L] = E[Q_0] + sum_l E[Q_l - Q_{l-1}] with the
    same inputs on both meshes of a level pair, and the variance telescoped
    the same way from unbiased sample variances, Var[Q_L] = Var[Q_0] +
    sum_l (Var[Q_l] - Var[Q_{l-1}]), each difference from the level's own
    paired samples (raw second moments would cancel catastrophically).  All
    level pairs advance together (one context per mesh and level, one stream
    each)."""
    t0 = time.perf_counter()
    meshes = {n: problem.mesh(n) for n in levels}
    groups, ins = [], []
    skip = 0
    for li, (n, N) in enumerate(zip(levels, samples)):
        inp = sample_inputs(N, problem.bounds, seed=seed, skip=skip)
        skip += N
        fine = make_batch(problem, meshes[n], inp)
        coarse = make_batch(problem, meshes[levels[li - 1]], inp) if li > 0 else None
        groups.append((fine, coarse))
        ins.append(inp)
    advance([s for g in groups for s in g if s is not None], log=log)
    mean = 0.0
    sec = 0.0
    se2 = 0.0
    detail = {"levels": []}
    for li, (fine, coarse) in enumerate(groups):
        qf = qoi(problem, fine)
        qc = qoi(problem, coarse) if coarse is not None else np.zeros_like(qf)
        y1 = qf - qc
        m1, v1 = moments(y1)
        _, vf = moments(qf)
        vc = moments(qc)[1] if coarse is not None else 0.0
        mean = mean + m1
        sec = sec + (vf - vc)
        se2 = se2 + v1 / y1.shape[0]
        detail["levels"].append({"n": levels[li], "samples": y1.shape[0],
                                 "var_y": float(np.mean(v1))})
        fine.close()
        if coarse is not None:
            coarse.close()
    return Estimate(mean, sec, np.sqrt(se2), time.perf_counter() - t0, detail)


# ---------------------------------------------------------------------------
# surrogate control-variate experiment (error per GPU-second)
# ---------------------------------------------------------------------------
def _unit(problem: Problem, inputs):
    return np.stack([inputs[b[0]] for b in problem.bounds], 1)


def _sync_time(t0):
    return time.perf_counter() - t0


def reference_chunk(problem: Problem, n_fine, start, count, seed=0, log=None, fine_batch=None):
    """Samples [start, start + count) of cv_experiment's plain-MC reference
    (the same scrambled Sobol points, seed + 1000): (QoI per sample, seconds),
    so that a reference longer than one sitting can be formed in parts."""
    t0 = time.perf_counter()
    _, Q = plain_mc(problem, n_fine, count, seed=seed + 1000, skip=start, batch=fine_batch, log=log)
    return Q, _sync_time(t0)


def cv_experiment(problem: Problem, n_fine, n_coarse, n_train, n_pilot, n_pair, n_surr, n_ref,
                  epochs=300, batch=64, seed=0, log=None, fine_batch=None, coarse_batch=None,
                  ref=None):
    """Surrogate control variate against plain MC on the same fine mesh.

    Reference: plain MC with n_ref fine samples (its own cost is not charged).
    Control variate: an MLP trained on n_train coarse-mesh solver samples, its
    coefficients fitted on n_pilot fine samples (n_pilot = 0: on the training
    samples, coarse solver against surrogate; either way independent of the
    paired samples, so the estimate stays unbiased), then n_pair fine samples
    and n_surr surrogate evaluations.  fine_batch / coarse_batch: samples per
    solver context on each mesh (all at once by default).  ref: (QoI of the
    reference's n_ref samples in order, their seconds) formed beforehand by
    reference_chunk, instead of solving them here.  Returns the
    realised mean and variance errors of both estimators against the reference
    (RMS over the QoIs with non-zero reference variance, each scaled by the
    reference's standard deviation), their GPU-seconds, and the GPU-seconds
    plain MC would need to reach the control variate's standard error."""
    from .surrogate import Surrogate, _alpha, control_variate
    say = log or (lambda *_: None)
    if ref is not None:
        Q_ref, t_ref = np.asarray(ref[0], dtype=np.float64), float(ref[1])
        if Q_ref.shape != (n_ref, problem.n_qoi):
            raise ValueError(f"reference of shape {Q_ref.shape}, expected ({n_ref}, {problem.n_qoi})")
    else:
        Q_ref, t_ref = reference_chunk(problem, n_fine, 0, n_ref, seed=seed, log=say,
                                       fine_batch=fine_batch)
    m_ref, v_ref = moments(Q_ref)
    keep = v_ref > 1e-24 * np.maximum(1.0, m_ref ** 2)
    sd = np.sqrt(v_ref[keep])
    per_fine_sample = t_ref / n_ref
    say(f"reference: {n_ref} samples at {n_fine}^2 in {t_ref:.1f} s")

    t0 = time.perf_counter()
    u_tr, Q_tr = plain_mc(problem, n_coarse, n_train, seed=seed + 1, batch=coarse_batch, log=say)
    t_train_sims = _sync_time(t0)
    t0 = time.perf_counter()
    sur = Surrogate(len(problem.bounds), problem.n_qoi, max_batch=max(batch, 4096), seed=seed + 2)
    losses = sur.fit(_unit(problem, u_tr), Q_tr, epochs=epochs, batch=batch)
    t_fit = _sync_time(t0)
    say(f"surrogate trained on {n_train} samples at {n_coarse}^2 in {t_fit:.1f} s, "
        f"final loss {float(losses[-1]) if len(losses) else float('nan'):.3g}")
    t0 = time.perf_counter()
    if n_pilot:
        u_a, Q_a = plain_mc(problem, n_fine, n_pilot, seed=seed + 3, batch=fine_batch, log=say)
    u_p, Q_p = plain_mc(problem, n_fine, n_pair, seed=seed + 4, batch=fine_batch, log=say)
    t_fine = _sync_time(t0)
    t0 = time.perf_counter()
    G_a = sur(_unit(problem, u_a if n_pilot else u_tr))
    G_p = sur(_unit(problem, u_p))
    G_m = sur(_unit(problem, sample_inputs(n_surr, problem.bounds, seed=seed + 5)))
    t_eval = _sync_time(t0)
    Q_fit = Q_a if n_pilot else Q_tr
    est = control_variate(Q_p, G_p, G_m, alpha=_alpha(Q_fit, G_a),
                          alpha2=_alpha(Q_fit ** 2, G_a ** 2))
    cv_seconds = t_train_sims + t_fit + t_fine + t_eval

    # plain MC with the same fine-sample budget as the control variate
    n_plain = n_pilot + n_pair
    Q_plain = np.vstack([Q_a, Q_p]) if n_pilot else Q_p
    m_pl, v_pl = moments(Q_plain)

    def err(m, v):
        return (float(np.sqrt(np.mean(((m[keep] - m_ref[keep]) / sd) ** 2))),
                float(np.sqrt(np.mean(((v[keep] - v_ref[keep]) / v_ref[keep]) ** 2))))

    e_cv, e_pl = err(est.mean, est.var), err(m_pl, v_pl)
    # GPU-seconds plain MC needs for the control variate's standard error
    se_ratio = np.median((np.sqrt(v_ref[keep] / n_pair) / np.maximum(est.mean_se[keep], 1e-300)) ** 2)
    plain_seconds_equal_error = float(se_ratio) * n_pair * per_fine_sample
    return {
        "reference": {"samples": n_ref, "mesh": n_fine, "seconds": t_ref},
        "control_variate": {"mean_error": e_cv[0], "var_error": e_cv[1], "gpu_seconds": cv_seconds,
                            "train_samples": n_train, "train_mesh": n_coarse, "pilot": n_pilot,
                            "paired": n_pair, "surrogate_evals": n_surr,
                            "final_loss": float(losses[-1]) if len(losses) else None,
                            "median_variance_reduction": float(np.median(est.variance_reduction[keep])),
                            "variance_reduction_integrals": [float(v) for v in est.variance_reduction[-2:]],
                            "alpha_fit": "pilot fine samples" if n_pilot else "training samples",
                            "seconds": {"train_sims": t_train_sims, "fit": t_fit, "fine": t_fine,
                                        "surrogate": t_eval}},
        "plain_mc": {"mean_error": e_pl[0], "var_error": e_pl[1], "samples": n_plain,
                     "gpu_seconds": n_plain * per_fine_sample},
        "plain_seconds_for_cv_error": plain_seconds_equal_error,
        "speedup_at_equal_error": plain_seconds_equal_error / cv_seconds,
        "qoi_kept": int(np.sum(keep)),
    }


def mlmc_experiment(problem: Problem, levels, samples, n_ref, seed=0, log=None, batch=None):
    """Multilevel MC against plain MC on the finest level at the same
    GPU-seconds, both measured against a plain-MC reference with n_ref samples
    on the finest mesh.  Errors: RMS over the QoIs with non-zero reference
    variance of |estimate - reference| scaled by the reference's standard
    deviation (means) or variance (variances).  batch: samples per context of
    the plain-MC runs on the finest mesh (all at once by default)."""
    say = log or (lambda *_: None)
    t0 = time.perf_counter()
    _, Q_ref = plain_mc(problem, levels[-1], n_ref, seed=seed + 1000, batch=batch, log=say)
    t_ref = time.perf_counter() - t0
    m_ref, v_ref = moments(Q_ref)
    keep = v_ref > 1e-24 * np.maximum(1.0, m_ref ** 2)
    say(f"reference: {n_ref} samples at {levels[-1]}^2 in {t_ref:.1f} s")
    est = mlmc(problem, levels, samples, seed=seed, log=say)
    say(f"multilevel estimate in {est.seconds:.1f} s")
    n_plain = max(2, int(round(est.seconds / (t_ref / n_ref))))
    t0 = time.perf_counter()
    _, Q_pl = plain_mc(problem, levels[-1], n_plain, seed=seed + 7, batch=batch, log=say)
    t_pl = time.perf_counter() - t0
    m_pl, v_pl = moments(Q_pl)

    def err(m, v):
        return (float(np.sqrt(np.mean(((m[keep] - m_ref[keep]) / np.sqrt(v_ref[keep])) ** 2))),
                float(np.sqrt(np.mean(((v[keep] - v_ref[keep]) / v_ref[keep]) ** 2))))

    e_ml, e_pl = err(est.mean, est.var), err(m_pl, v_pl)
    sd = np.sqrt(v_ref[keep])
    return {"reference": {"samples": n_ref, "mesh": levels[-1], "seconds": t_ref,
                          "mean_se": float(np.sqrt(1.0 / n_ref))},
            "mlmc": {"levels": list(levels), "samples": list(samples), "seconds": est.seconds,
                     "mean_error": e_ml[0], "var_error": e_ml[1],
                     "mean_se": float(np.sqrt(np.mean((est.mean_se[keep] / sd) ** 2))),
                     "detail": est.detail},
            "plain_mc": {"samples": n_plain, "seconds": t_pl, "mean_error": e_pl[0],
                         "var_error": e_pl[1], "mean_se": float(np.sqrt(1.0 / n_plain))},
            "qoi_kept": int(np.sum(keep))}
