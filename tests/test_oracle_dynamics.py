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


# -- the Cash-Karp step itself, on scalar problems with known solutions -----

_DERIVS = None


def _rkck(f, y0, x, h):
    """One or_rkck step of y' = f(t, y) (scalar) from (x, y0): (y5, yerr)."""
    import ctypes
    global _DERIVS
    lib = O.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    if _DERIVS is None:
        _DERIVS = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, dp, dp)
    cb = _DERIVS(lambda ctx, t, q, dq: dq.__setitem__(0, f(t, q[0])))
    lib.or_rkck.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                            dp, dp, _DERIVS]
    lib.or_rkck.restype = None
    y = np.array([y0] + [0.0] * 8)
    dydx = np.array([f(x, y0)] + [0.0] * 8)
    yout = np.zeros(9)
    yerr = np.zeros(9)
    lib.or_rkck(None, y.ctypes.data_as(dp), dydx.ctypes.data_as(dp), 1, x, h,
                yout.ctypes.data_as(dp), yerr.ctypes.data_as(dp), cb)
    return yout[0], yerr[0]


def test_rkck_fifth_order_and_embedded_fourth():
    """y' = -2 t y^2, y(0) = 1 (y = 1 / (1 + t^2)): nonlinear and
    non-autonomous, so a wrong Butcher coefficient or node lowers the
    order.  Halving h divides the local error of the fifth-order solution
    by about 2^6 and the embedded error estimate (fifth minus fourth order)
    by about 2^5."""
    f = lambda t, y: -2.0 * t * y * y
    exact = lambda t: 1.0 / (1.0 + t * t)
    x = 0.3
    e5, ee = [], []
    for h in (0.2, 0.1, 0.05):
        y5, yerr = _rkck(f, exact(x), x, h)
        e5.append(abs(y5 - exact(x + h)))
        ee.append(abs(yerr))
    for a, b in zip(e5, e5[1:]):
        assert 40.0 < a / b < 100.0, e5
    for a, b in zip(ee, ee[1:]):
        assert 22.0 < a / b < 45.0, ee


def test_rkck_keeps_the_single_precision_node():
    """y' = t: every order condition of the fifth-order weights holds, so the
    step is exact but for the node a4, which the reference writes as the
    REAL*4 literal .6 (suppl_routines.f:1301): the increment exceeds h^2/2
    by C4 h^2 (fl32(0.6) - 0.6), C4 = 125/594."""
    x, h = 0.0, 1.0
    y5, _ = _rkck(lambda t, y: t, 0.0, x, h)
    shift = 125.0 / 594.0 * h * h * (float(np.float32(0.6)) - 0.6)
    assert abs(shift) > 4e-9
    assert abs(y5 - (h * h / 2.0 + shift)) < 1e-15
