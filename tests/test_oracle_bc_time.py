"""Pins of the oracle's boundary values of time.  The reference evaluates a
boundary's expressions at (x, y, int_step_time) and sets
int_step_time = elapsed_time + r*dt_global before stage r only in Heun's
scheme (src/fv.cc:2296-2298, src/bc.h:411-562); its own Heun update gives NaN
after one step, so nothing of its output can be matched.
 * The callback is called before each stage with the stage times that the
   run's own log implies: t^n, then t^n + dt.
 * One step equals the two-stage SSP scheme built by hand from the oracle's
   residual with the boundary values set at t^n for the first stage and at
   t^n + dt for the second.
 * With boundary values that vary smoothly in time the scheme stays second
   order in time; the same test with the second stage's values left at t^n
   (the plausible mistake) measures first order.
"""
import numpy as np

import tensum_amd as T
import oracle_binding as O
from cases import conserved

# inlet (tag 4), farfield (3), pressure outlet (2), noslip wall (1)
ST = (1.0, 0.4, 0.05, 1.0)
BC = {4: ("inlet", ST), 3: ("farfield", ST), 2: ("pressure", ST),
      1: ("noslip", (1.2, 0.0, 0.0, 1.0))}

# field expressions (T, u, v, p) of x, y, t per tag; None keeps the table value
EXPR = {
    4: ("1.0 + 0.1*sin(4*pi*t + y)", "0.4 + 0.1*cos(3*t)*y", None, "1.0 + 0.05*t*(x + y)"),
    3: (None, "0.4 + 0.2*t", "0.05*exp(-t)", "1 + 0.1*sin(5*t)^2"),
    2: (None, None, None, "1.0 - 0.1*tanh(3*t)"),
    1: ("1.2 + 0.1*t", "0.05*sin(6*t)", None, None),
}


def _fields(tag, x, y, t):
    """numpy restatement of EXPR (the values the device programs must give)."""
    pi = np.pi
    if tag == 4:
        return (1.0 + 0.1 * np.sin(4 * pi * t + y), 0.4 + 0.1 * np.cos(3 * t) * y, None,
                1.0 + 0.05 * t * (x + y))
    if tag == 3:
        return None, 0.4 + 0.2 * t, 0.05 * np.exp(-t), 1 + 0.1 * np.sin(5 * t) ** 2
    if tag == 2:
        return None, None, None, 1.0 - 0.1 * np.tanh(3 * t)
    return 1.2 + 0.1 * t, 0.05 * np.sin(6 * t), None, None


def values_at(m, t):
    """(n_bface, 2, 4) boundary values at time t: table values, replaced by
    the EXPR fields where given."""
    vals = np.zeros((m.bface_tag.size, 2, 4))
    bv = m.bface_v.reshape(-1, 2)
    for b, tag in enumerate(m.bface_tag):
        vals[b, :, :] = m.bc[int(tag)][1]
    for e in range(2):
        x, y = m.x[bv[:, e]], m.y[bv[:, e]]
        for tag in EXPR:
            sel = m.bface_tag == tag
            f = _fields(tag, x[sel], y[sel], t)
            for q in range(4):
                if f[q] is not None:
                    vals[sel, e, q] = f[q]
    return vals


def callback(m, mistake=False):
    """Oracle callback writing values_at(t); with mistake=True the second
    stage of each step gets the first stage's time."""
    calls = []

    def fn(t, a):
        calls.append(t)
        tt = calls[-2] if (mistake and len(calls) % 2 == 0) else t
        a[...] = values_at(m, tt)
    return fn, calls


def state(m):
    return np.stack([1.0 + 0.1 * np.sin(np.pi * m.x), 0.4 + 0.1 * m.y,
                     0.05 * np.cos(np.pi * m.x), 1.0 + 0.05 * np.sin(np.pi * m.y)], 1)


def prm(**kw):
    return T.Params(**{**dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                              final_time=10.0, time_scheme="heuns"), **kw})


def test_callback_gets_each_stage_time():
    m = T.jittered_mesh(17, seed=5, bc=BC)
    o = O.Oracle(m, prm())
    o.set_state(state(m))
    fn, calls = callback(m)
    o.set_bc_time_fn(fn)
    n, log = o.run(6, want_log=True)
    assert n == 6 and len(calls) == 12
    t_prev = np.concatenate([[0.0], log[:-1, 2]])
    assert np.array_equal(calls[0::2], t_prev)
    assert np.array_equal(calls[1::2], t_prev + log[:, 1])
    assert np.array_equal(calls[1::2], log[:, 2])  # elapsed + dt is the next t^n


def test_one_step_is_the_two_stage_scheme_with_stage_boundary_values():
    m = T.jittered_mesh(17, seed=6, bc=BC)
    W0 = state(m)
    o = O.Oracle(m, prm())
    t0 = 0.3
    o.set_bc_state(values_at(m, t0))
    o.set_state(W0)
    dt = o.compute_dt()
    A = o.geometry()["dcarea"][:, None]

    def L(W, t):
        o.set_bc_state(values_at(m, t))
        o.set_state(W)
        return -(dt / A) * o.residual()

    def prim(U, gam=1.4, R=1.0):
        rho, mx, my, E = U.T
        u, v = mx / rho, my / rho
        p = (gam - 1) * (E - 0.5 * rho * (u * u + v * v))
        return np.stack([p / (R * rho), u, v, p], 1)

    U0 = conserved(W0)
    U1 = U0 + L(W0, t0)
    U2 = 0.5 * U0 + 0.5 * (U1 + L(prim(U1), t0 + dt))
    s = O.Oracle(m, prm())
    s.set_state(W0)
    s.set_clock(0, 1, t0, 0.0)
    fn, calls = callback(m)
    s.set_bc_time_fn(fn)
    assert s.run(1) == 1
    assert calls == [t0, t0 + dt]
    np.testing.assert_allclose(s.get_state(), prim(U2), rtol=1e-12, atol=1e-14)


def _converge(mistake):
    m = T.jittered_mesh(13, seed=7, bc=BC)
    W0 = state(m)

    def run(dt, t=0.2):
        o = O.Oracle(m, prm(cfl=0.0, time_step=dt))
        o.set_state(W0)
        fn, _ = callback(m, mistake)
        o.set_bc_time_fn(fn)
        assert o.run(int(round(t / dt))) == int(round(t / dt))
        return conserved(o.get_state())

    ref = run(0.2 / 512)
    errs = [np.abs(run(0.2 / k) - ref).max() for k in (16, 32, 64)]
    return np.log2(np.array(errs[:-1]) / np.array(errs[1:])), errs


def test_second_order_in_time_with_boundary_values_of_time():
    rates, errs = _converge(False)
    assert np.all(rates > 1.8) and np.all(rates < 2.3), (errs, rates)


def test_second_stage_at_the_first_stage_time_is_first_order():
    rates, errs = _converge(True)
    assert np.all(rates < 1.3), (errs, rates)
