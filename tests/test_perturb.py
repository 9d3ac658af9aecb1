"""pert(type, x, y, p) of param.in expressions (src/perturb.cc) against
hand-computed values."""
import math

import pytest

import tensum_amd as T


def test_pert4_picks_the_pth_number():
    assert T.pert(4, 0.0, 0.0, 2, [0.3, 0.7]) == 0.7
    assert T.pert(4, 9.0, 9.0, 1, [0.3, 0.7]) == 0.3
    assert T.pert(4, 0.0, 0.0, 5, []) == 0.0  # no numbers: zero, as the reference
    with pytest.raises(T.TsmError):
        T.pert(4, 0.0, 0.0, 3, [0.3, 0.7])


def test_pert1_and_pert3_multimodal_band():
    r = [1.0, 1.0, 0.0, math.pi / 2]  # two modes, equal weights, phases 0 and pi/2
    assert T.pert(1, 0.0, 0.5, 1.0, r) == pytest.approx(-0.5, abs=1e-15)
    assert T.pert(1, 0.0, 0.25, 1.0, r) == pytest.approx(0.5, abs=1e-15)
    assert T.pert(1, 0.0, 0.25, 0.2, r) == 0.0          # outside the band (0.4, 0.6)
    assert T.pert(3, 0.0, 0.5, 1.0, r) == pytest.approx(-0.5, abs=1e-15)  # mode 1: cos(2 pi y)
    assert T.pert(3, 0.0, 0.25, 1.0, r) == pytest.approx(0.0, abs=1e-15)


def test_pert2_piecewise_constant_with_alternating_sign():
    r = [0.1, 0.2, 0.3, 0.4]
    assert [T.pert(2, 0.0, y, 1.0, r) for y in (0.1, 0.3, 0.6, 0.9)] == [0.1, -0.2, 0.3, -0.4]
    assert T.pert(2, 0.0, 0.3, 0.5, r) == 0.1           # band (0.25, 0.75), cells of 0.125
    assert T.pert(2, 0.0, 0.8, 0.5, r) == 0.0


def test_unknown_type_is_an_error():
    with pytest.raises(T.TsmError):
        T.pert(5, 0.0, 0.0, 1.0, [1.0])
