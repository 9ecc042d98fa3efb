"""GPU parity of the batched CC-MQCT form: mqct_hip_propagate with a CC mode
and batched >= 1 runs every trajectory's classical and quantum equations
(DERIVS_BK, mqct_eqtns.f:112-499) and DOTRAJECT's dt intervals
(propagator.f:3873-4050) in the interval-bucketed form, whose GEMM also
forms dM/dR v for the mean-field force and the Coriolis sums.

Against the CPU oracle's DOTRAJECT restatement (diatom + atom, Theta
equations off and on) and against the plain per-trajectory kernel
(asymmetric top + atom, every GEMM shape, FP32 coefficient forms).  Bars
as for the plain form: fixed-step RK4, amplitudes within 1e-9 and
classical variables within 1e-9 relative, the same step counts; odeint,
1e-6 (its step-size decisions may round differently; the tolerance is
1e-8 per step); FP32 knot matrices, 1e-6 of the FP64 form.
"""
import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.basis import CM_TO_HARTREE, AMUTOAU
from mqct_amd.propagate import Propagator
import oracle.oracle as O

pytestmark = pytest.mark.gpu

U_CM = 500.0
N_SHAPES = 23


@pytest.fixture(scope="module")
def c0():
    wl = W.diatom_atom()
    orc = O.oracle_for_workload(wl)
    M, D, S = orc.matrix_make(wl.basis.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    return wl, M, D


def compare(q_gpu, q_ora, N, amp_tol, cl_tol):
    a = np.max(np.abs(q_gpu[:2 * N] - q_ora[:2 * N]))
    c = np.max(np.abs(q_gpu[2 * N:2 * N + 6] - q_ora[2 * N:2 * N + 6]) /
               np.maximum(np.abs(q_ora[2 * N:2 * N + 6]), 1e-6))
    assert a <= amp_tol, f"amplitude diff {a}"
    assert c <= cl_tol, f"classical diff {c}"
    return a, c


@pytest.mark.parametrize("theta_on", [0, 1])
@pytest.mark.parametrize("mode", [L.MQCT_PROP_CC_RK4, L.MQCT_PROP_CC_ODEINT])
def test_cc_batched_matches_oracle(dev, c0, mode, theta_on):
    """Six trajectories from L = 0 to the largest L of the grid, all in one
    batch, against the oracle one by one; the plain kernel is held to the
    same bars on the same rows.  odeint with the Theta equations on leaves
    out L = 11 and 23: there P_Theta stays at rounding level (1e-16, from
    cos(pi/2) != 0) and odeint's relative error control, floored only by
    TINY = 1e-30, chases it down to MIN_TMSTP -- in the oracle and the
    plain kernel as much as here (over 20000 steps, with results that
    depend on the rounding)."""
    wl, M, D = c0
    b = wl.basis
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, M, D, wl.mass_amu * AMUTOAU,
                      theta_on=theta_on)
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu, theta_on=theta_on)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    pick = [0, 47, 71, 82] if (theta_on and mode == L.MQCT_PROP_CC_ODEINT) else [0, 11, 23, 47, 71, 82]
    rows = prop.trajectories([ls[i] for i in pick], U_CM, s_st=1, time_step=10.0, eps=1e-8, hmin=1e-4)
    q, st = prop.run(mode, rows, batched=1)
    qp, sp = prop.run(mode, rows, batched=0)
    N = b.n_states
    tol = (1e-9, 1e-9) if mode == L.MQCT_PROP_CC_RK4 else (1e-6, 1e-6)
    worst = 0.0
    for i, r in enumerate(rows):
        qo, so, _ = dyn.dotraject(mode, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"], hmin=r["hmin"],
                                  R_start=r["R_start"])
        assert st["fail"][i] == so["fail"] == 0
        if mode == L.MQCT_PROP_CC_RK4:
            assert st["n_steps"][i] == so["n_steps"] == sp["n_steps"][i]
            assert st["n_rhs"][i] == so["n_rhs"]
        assert abs(st["t_end"][i] - so["t_end"]) <= 1e-9 * so["t_end"]
        a, c = compare(q[i], qo, N, *tol)
        compare(qp[i], qo, N, *tol)
        worst = max(worst, a, c)
        assert st["norm_err"][i] < 1e-6
    print(f"mode {mode} theta_on {theta_on}: batched vs oracle worst {worst:.2e}")


def _asym_rows(dev, n, seed, eps=1e-8, hmin=1e-6):
    from mqct_amd.matrix import MatrixBuilder
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=120)
    bb, E = W.at_batch(n, seed=seed)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    k = np.sqrt(prop.mass * 2.0 * np.asarray(E) * CM_TO_HARTREE)
    l = np.floor(np.asarray(bb) * k)
    rows = []
    for i in range(n):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, eps=eps, hmin=hmin)
    return wl, prop, rows


@pytest.mark.parametrize("mode", [L.MQCT_PROP_CC_RK4, L.MQCT_PROP_CC_ODEINT])
def test_cc_batched_matches_plain_asym_top(dev, mode):
    """Asymmetric top + atom (j <= 3: 84 states in 7 m-blocks), 32
    trajectories of uneven length (b 0-18 bohr, E 200-2000 cm^-1):
    batched against the plain kernel.  MIN_TMSTP = 20 a.u. for odeint, as
    in the AT tests (amplitudes at rounding-noise level otherwise drive
    the relative error control to the minimum step)."""
    wl, prop, rows = _asym_rows(dev, 32, seed=17, hmin=20.0 if mode == L.MQCT_PROP_CC_ODEINT else 1e-6)
    q0, s0 = prop.run(mode, rows, batched=0)
    q1, s1 = prop.run(mode, rows, batched=1)
    N = wl.basis.n_states
    tol = 1e-9 if mode == L.MQCT_PROP_CC_RK4 else 1e-6
    for i in range(len(rows)):
        compare(q1[i], q0[i], N, tol, tol)
    assert np.array_equal(s0["n_steps"], s1["n_steps"])
    assert np.array_equal(s0["fail"], s1["fail"])
    assert np.allclose(s0["t_end"], s1["t_end"], rtol=1e-12)
    assert np.all(s1["norm_err"] < 1e-6)


@pytest.mark.parametrize("shape", list(range(N_SHAPES)))
def test_cc_batched_every_gemm_tile(dev, shape):
    """Every interval-GEMM shape with the CC entries (four products per
    trajectory, two bucket entries per knot): the plain kernel's RK4
    trajectories (asymmetric top + atom, j <= 3, 24 trajectories)."""
    wl, prop, rows = _asym_rows(dev, 24, seed=23)
    q0, s0 = prop.run(L.MQCT_PROP_CC_RK4, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
    try:
        q1, s1 = prop.run(L.MQCT_PROP_CC_RK4, rows, batched=1)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    for i in range(len(rows)):
        compare(q1[i], q0[i], N, 1e-9, 1e-9)
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


@pytest.mark.parametrize("batched", [2, 3])
def test_cc_batched_fp32_coefficients(dev, batched):
    """FP32 knot matrices (batched=2) and FP32 products in the pipelined
    forms (batched=3): a fixed perturbation of the coupling and of its R
    derivative at FP32 rounding, so the trajectories stay within 1e-6 of
    the FP64-coefficient form with the same step counts."""
    wl, prop, rows = _asym_rows(dev, 24, seed=29)
    q64, s64 = prop.run(L.MQCT_PROP_CC_RK4, rows, batched=1)
    q32, s32 = prop.run(L.MQCT_PROP_CC_RK4, rows, batched=batched)
    N = wl.basis.n_states
    d = 0.0
    for i in range(len(rows)):
        d = max(d, max(compare(q32[i], q64[i], N, 1e-6, 1e-6)))
    assert d > 0.0
    assert np.array_equal(s32["n_steps"], s64["n_steps"])
    assert np.all(s32["norm_err"] < 1e-6)


def test_cc_batched_step_cap(dev, c0):
    """max_steps ends a trajectory with fail = 2 after that many steps, as
    the plain form does (host-forced: a small cap on a normal run)."""
    wl, M, D = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    rows = prop.trajectories([ls[5], ls[40]], U_CM, s_st=1, time_step=10.0, max_steps=25)
    q0, s0 = prop.run(L.MQCT_PROP_CC_RK4, rows, batched=0)
    q1, s1 = prop.run(L.MQCT_PROP_CC_RK4, rows, batched=1)
    assert list(s1["fail"]) == list(s0["fail"]) == [2, 2]
    assert list(s1["n_steps"]) == list(s0["n_steps"]) == [25, 25]
    N = b.n_states
    for i in range(2):
        compare(q1[i], q0[i], N, 1e-9, 1e-9)
