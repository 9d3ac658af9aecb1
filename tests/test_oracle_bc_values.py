"""Pins of the oracle's per-end-point boundary values (the reference evaluates a
boundary's expressions at each end point's coordinates with the sample's
random numbers, src/bc.h:411-562).
 * Values equal to the tag's constant state give the constant-state run bit
   for bit.
 * The values belong to vertices, not to the order a bface lists them in:
   reversing every bface's vertex order in the input, with its values swapped
   along, gives the same run.
 * They take effect: values that differ from the table change the run, and a
   boundary value at one end point changes only residuals near that end point.
"""
import numpy as np

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2

ST = (1.0, 0.4, 0.05, 1.0)
BC = {4: ("inlet", ST), 3: ("farfield", ST), 2: ("pressure", ST),
      1: ("noslip", (1.2, 0.0, 0.0, 1.0))}


def _prm(**kw):
    return T.Params(**{**dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                              final_time=10.0), **kw})


def _state(m):
    return np.stack([1.0 + 0.1 * np.sin(np.pi * m.x), 0.4 + 0.1 * m.y, 0.05 * np.cos(np.pi * m.x),
                     1.0 + 0.05 * np.sin(np.pi * m.y)], 1)


def table_values(m):
    vals = np.zeros((m.bface_tag.size, 2, 4))
    for b, tag in enumerate(m.bface_tag):
        vals[b, :, :] = m.bc[int(tag)][1]
    return vals


def smooth_values(m, amp=0.1):
    """Table values times a smooth function of each end point's position."""
    vals = table_values(m)
    bv = m.bface_v.reshape(-1, 2)
    for e in range(2):
        x, y = m.x[bv[:, e]], m.y[bv[:, e]]
        f = 1.0 + amp * np.sin(3 * x + 2 * y)
        vals[:, e, 0] *= f
        vals[:, e, 1] *= f
        vals[:, e, 3] *= 2.0 - f
    return vals


def _run(m, prm, vals=None, steps=20):
    o = O.Oracle(m, prm)
    if vals is not None:
        o.set_bc_state(vals)
    o.set_state(_state(m))
    R = o.residual()
    o.run(steps)
    return R, o.get_state()


def test_table_values_give_the_constant_state_run():
    m = T.jittered_mesh(17, seed=2, bc=BC)
    for prm in (_prm(), _prm(model="ns", mu_ref=0.01, flux="kepec", reconstruct="first")):
        a = _run(m, prm)
        b = _run(m, prm, table_values(m))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_values_follow_vertices_not_bface_order():
    m = T.jittered_mesh(17, seed=2, bc=BC)
    vals = smooth_values(m)
    flipped = T.Mesh(m.x, m.y, m.bface_tag, m.bface_v.reshape(-1, 2)[:, ::-1].ravel().copy(),
                     m.cell_v, m.periodic_ptr, m.periodic_v, m.bc)
    for prm in (_prm(), _prm(model="ns", mu_ref=0.01, flux="kepec", reconstruct="first")):
        a = _run(m, prm, vals)
        b = _run(flipped, prm, vals[:, ::-1, :].copy())
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_values_take_effect_locally():
    m = T.jittered_mesh(17, seed=2, bc=BC)
    base = table_values(m)
    R0, W0 = _run(m, _prm(), base)
    assert rel_l2(conserved(_run(m, _prm(), smooth_values(m))[1]), conserved(W0)) > 1e-4
    # one end point of one inlet bface gets a hotter inlet state
    b = int(np.flatnonzero(m.bface_tag == 4)[3])
    v = int(m.bface_v[2 * b])
    one = base.copy()
    one[b, 0, 0] *= 1.5
    R1, _ = _run(m, _prm(), one, steps=0)
    changed = np.flatnonzero(np.abs(R1 - R0).max(1) > 0)
    assert v in changed
    # only the vertex itself and vertices of its cells and their neighbours
    # (gradient-coupled) see the change
    d = np.hypot(m.x[changed] - m.x[v], m.y[changed] - m.y[v])
    assert d.max() < 3.5 / 16
