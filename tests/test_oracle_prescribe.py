"""Known answers for prescribed motion (PRESCRIBE, input.cpp:2647-2795;
prescribe_acceleration / prescribe_velocity, input.cpp:4247-4421) on the
oracle, and the deck API's tabulation of its histories."""
import math

import numpy as np
import pytest

from parmec_amd import pmx
from test_oracle_kat import _one_sphere, _oracle_run


def _prescribed(m, lin_kind, ang_kind):
    m.set_prescribed(1)
    m.prspart[0] = 0
    m.prskind[0, 0] = lin_kind
    m.prskind[1, 0] = ang_kind
    return m


def test_prescribed_velocity_drift():
    """Linear velocity prescribed to u from the first step: the force is
    zeroed (gravity ignored), so after K steps x = x0 + 1.5 h v0 + (K-1) h u
    (half step of the first call, the first step still at v0)."""
    h, K = 1e-3, 200
    v0, u = np.array([0.3, 0.0, -0.1]), np.array([1.0, -2.0, 0.5])
    m = _prescribed(_one_sphere(v=v0), 0, -1)
    tab = np.zeros((K, 1, 6))
    tab[:, 0, :3] = u
    mm, o = _oracle_run(m, (K - 0.5) * h, h, gravity=(0.0, 0.0, -9.81), prstab=tab)
    assert o.stats()["steps"] == K
    np.testing.assert_array_equal(mm.linear[:, 0], u)
    np.testing.assert_allclose(mm.position[:3, 0], 1.5 * h * v0 + (K - 1) * h * u, rtol=1e-12, atol=1e-13)


def test_prescribed_acceleration_is_free_flight():
    """Linear acceleration prescribed to a: f = m a replaces gravity and
    contacts, so the particle follows the leapfrog free-flight closed form
    with g = a."""
    h, K = 1e-3, 300
    v0, a = np.array([1.0, -2.0, 3.0]), np.array([0.0, 4.0, -1.5])
    m = _prescribed(_one_sphere(v=v0), 1, -1)
    tab = np.zeros((K, 1, 6))
    tab[:, 0, :3] = a
    mm, o = _oracle_run(m, (K - 0.5) * h, h, gravity=(0.0, 0.0, -9.81), prstab=tab)
    x = (K + 0.5) * h * v0 + h * h * a * K * (K + 1) / 2.0
    np.testing.assert_allclose(mm.position[:3, 0], x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(mm.linear[:, 0], v0 + K * h * a, rtol=1e-12, atol=1e-12)


def test_prescribed_angular_velocity_sets_body_rate():
    """Angular velocity prescribed to w (spatial): omega = w and the body
    rate is L^T w with the rotation of the step; a spin about z turns the
    body about z."""
    h, K, wz = 1e-3, 400, 2.0
    m = _prescribed(_one_sphere(), -1, 0)
    tab = np.zeros((K, 1, 6))
    tab[:, 0, 5] = wz
    mm, o = _oracle_run(m, (K - 0.5) * h, h, prstab=tab)
    np.testing.assert_array_equal(mm.angular[3:, 0], [0.0, 0.0, wz])
    L = mm.rotation[:, 0].reshape(3, 3, order="F")
    np.testing.assert_allclose(mm.angular[:3, 0], L.T @ np.array([0.0, 0.0, wz]), atol=1e-15)
    ang = (K - 1) * h * wz
    Rz = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    np.testing.assert_allclose(L, Rz, atol=1e-9)


def test_deck_prescribe_tabulates_histories():
    """PRESCRIBE with a TSERIES tuple and a callable: kinds and the per-step
    values reach the engine (checked on the oracle through the deck)."""
    from oracle import orc
    from parmec_amd import parmec as P
    P.RESET()
    mat = P.MATERIAL(1000.0, 1e6, 0.25)
    P.SPHERE((0.0, 0.0, 0.0), 0.1, mat, 1)
    P.SPHERE((1.0, 0.0, 0.0), 0.1, mat, 1)
    ts = [P.TSERIES(0.5), P.TSERIES([0.0, 0.0, 1.0, 1.0]), P.TSERIES(-1.0)]
    P.PRESCRIBE(0, linear=tuple(ts), kind="vv")
    P.PRESCRIBE(1, angular=lambda t: (0.0, 0.0, 3.0 * t), kind="va")
    with pytest.raises(ValueError):
        P.PRESCRIBE(0, linear=(0, 1), kind="vv")
    with pytest.raises(ValueError):
        P.PRESCRIBE(0, linear=tuple(ts), kind="xv")
    P._use_engine(orc.Oracle(1))
    P.DEM(0.1, 1e-3)
    m = P.state()
    t_last = m.curtime - 1e-3
    np.testing.assert_allclose(m.linear[:, 0], [0.5, t_last, -1.0], rtol=1e-12)
    assert list(m.prskind[:, 0]) == [0, -1] and list(m.prskind[:, 1]) == [-1, 1]
    P.RESET()
