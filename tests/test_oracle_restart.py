"""Pins of restart continuation (src/fv.cc:300-402, :2230-2245): a run stopped
after k steps and continued from its state and its loop clock (iteration,
stamp counter, time, first-step residual norm: the four numbers a restart
file carries after the state, src/writer.cc:404-407) is the uninterrupted run,
bit for bit, in unsteady mode with time stamps and in steady mode (where the
first-step norm normalises every later residual).  Without the clock the
continuation differs.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import vortex, vortex_mesh

CASES = {
    "unsteady_stamps": T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=0.6,
                                time_stamps=4),
    "steady": T.Params(flux="kepec", reconstruct="first", cfl=0.8, time_mode="steady", max_iter=40,
                       min_residue=1e-12),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_continuation_from_state_and_clock_is_the_uninterrupted_run(case):
    prm = CASES[case]
    m = vortex_mesh(17)
    a = O.Oracle(m, prm)
    a.set_state(vortex(m.x, m.y))
    n, la = a.run(10 ** 4, want_log=True)
    assert n > 12
    for k in (1, 5, n - 3):
        b = O.Oracle(m, prm)
        b.set_state(vortex(m.x, m.y))
        assert b.run(k) == k
        clock = b.clock()
        assert clock[0] == k and clock[2] == la[k - 1, 2] and clock[3] > 0
        c = O.Oracle(m, prm)
        c.set_state(b.get_state())
        c.set_clock(*clock)
        nc, lc = c.run(10 ** 4, want_log=True)
        assert nc == n - k
        assert np.array_equal(lc, la[k:])
        assert np.array_equal(c.get_state(), a.get_state())
        assert c.clock() == a.clock()
        if k == 5:  # the state alone restarts the clock
            d = O.Oracle(m, prm)
            d.set_state(b.get_state())
            nd, ld = d.run(n - k, want_log=True)
            assert not np.array_equal(ld[:, 2:], la[k:, 2:])
