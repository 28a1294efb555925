"""Host-side logic of the mirror that needs no GPU: the BDF/EXT coefficient
ramp and the TimeStepper history (time.jl:31-112) against the oracle and the
textbook closed forms."""
import numpy as np
import pytest

from oracle import sem_oracle as o


@pytest.fixture(scope="module")
def cd():
    try:
        from semhip import cd as m
    except OSError as e:  # library missing
        pytest.skip(f"libsemhip.so not loadable: {e}")
    return m


@pytest.mark.parametrize("t", [[0.0] * 4, [1e-3, 0, 0, 0], [2e-3, 1e-3, 0, 0],
                               [3e-3, 2e-3, 1e-3, 0], [0.7, 0.5, 0.2, 0.1]])
def test_bdfextk_matches_oracle_bitwise(cd, t):
    for k in (1, 2, 3):
        a, b = cd.bdfExtK(t, k=k)
        ao, bo = o.bdfExtK(t, k=k)
        np.testing.assert_array_equal(a, ao)
        np.testing.assert_array_equal(b, bo)


def test_timestepper_ramps_bdf1_bdf2(cd):
    dt = 1e-3
    ts = cd.TimeStepper(0.0, 1.0, dt, k=2)
    np.testing.assert_array_equal(ts.bdfB, [0, 0, 0])   # steady start
    ts.updateHist_()
    np.testing.assert_allclose(ts.bdfB, [1 / dt, -1 / dt, 0], rtol=1e-12)
    np.testing.assert_allclose(ts.bdfA, [1, 0], rtol=1e-12)
    for _ in range(3):
        ts.updateHist_()
        np.testing.assert_allclose(ts.bdfB, [1.5 / dt, -2 / dt, 0.5 / dt], rtol=1e-9)
        np.testing.assert_allclose(ts.bdfA, [2, -1], rtol=1e-9, atol=1e-12)
    st = ts.state()
    ts.updateHist_()
    ts.restore(st)
    assert ts.istep == 4 and ts.time == st[0]
