"""Pins of the oracle's propagation half (DERIVS_BK_adia, rk4, odeint) by
closed forms: a detuned two-level system with constant coupling (Rabi
formula), a purely diagonal coupling (pure phase), and the convergence
order of rk4."""
import numpy as np

import oracle.oracle as O


def two_level(c, e1, e2, m11=0.0, m22=0.0):
    """Two m=0 states of two channels, coupling c constant in R."""
    R = np.linspace(3.0, 12.0, 10)
    ind = np.array([[1, 1], [2, 1], [2, 2]], np.int32)
    mat = np.array([np.full(10, m11), np.full(10, c), np.full(10, m22)])
    der = np.zeros_like(mat)
    dyn = O.DynOracle(np.array([0, 0]), np.array([0, 0]), np.array([e1, e2]), ind, R, mat, der, mass=1000.0)
    t = np.linspace(0.0, 400.0, 41)
    path = np.zeros((41, 8))
    path[:, 0] = t
    path[:, 1] = 5.0              # R
    path[:, 3] = np.pi / 2        # Theta
    # P_Phi = 0: no Coriolis coupling
    dyn.set_path(path)
    return dyn, t[-1]


def test_rabi_detuned_closed_form():
    c, e1, e2 = 0.01, 0.0, 0.012
    dyn, T = two_level(c, e1, e2)
    q, st, _ = dyn.dotraject(2, 1, 0.0, 1.0, 0.05)  # AT + rk4
    d = (e2 - e1) / 2.0
    Om = np.sqrt(c * c + d * d)
    P2 = c * c / Om ** 2 * np.sin(Om * T) ** 2
    N = 2
    p = q[:N] ** 2 + q[N:2 * N] ** 2
    assert abs(p[1] - P2) < 1e-9
    assert abs(p.sum() - 1.0) < 1e-12
    # phase of a_1 in the interaction picture: a1 = e^{-i d t}(cos Om t + i d/Om sin Om t)
    a1 = q[0] + 1j * q[N]
    ref = np.exp(-1j * d * T) * (np.cos(Om * T) + 1j * d / Om * np.sin(Om * T))
    assert abs(a1 - ref) < 1e-9


def test_diagonal_coupling_is_a_pure_phase():
    dyn, T = two_level(0.0, 0.0, 0.0, m11=0.003, m22=-0.002)
    q, st, _ = dyn.dotraject(3, 1, 0.0, 1.0, 1.0, eps=1e-12, hmin=1e-8)  # AT + odeint
    a1 = q[0] + 1j * q[2]
    assert abs(a1 - np.exp(-1j * 0.003 * T)) < 1e-9
    assert abs(q[1]) + abs(q[3]) < 1e-15


def test_rk4_is_fourth_order():
    c, e1, e2 = 0.02, 0.0, 0.03
    d = (e2 - e1) / 2.0
    Om = np.sqrt(c * c + d * d)
    errs = []
    for dt in (4.0, 2.0, 1.0):
        dyn, T = two_level(c, e1, e2)
        q, st, _ = dyn.dotraject(2, 1, 0.0, 1.0, dt)
        a1 = q[0] + 1j * q[2]
        ref = np.exp(-1j * d * T) * (np.cos(Om * T) + 1j * d / Om * np.sin(Om * T))
        errs.append(abs(a1 - ref))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12.0 < r1 < 20.0 and 12.0 < r2 < 20.0
