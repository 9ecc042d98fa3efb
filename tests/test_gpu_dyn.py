"""GPU parity of the propagation (the HIP library through its C ABI)
against the CPU oracle's DOTRAJECT restatement.

The matrix (Mat_el, Mat_el_der) is the oracle's FP64 build of the
diatom+atom workload, so both sides integrate the same equations.  Bars:
fixed-step RK4 (same step sequence on both sides): amplitudes within 1e-9
absolute, classical variables within 1e-9 relative.  Adaptive odeint: the
step-size decisions may round differently, so amplitudes within 1e-6
(the tolerance is 1e-8 per step).  Norm drift is reported.
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


@pytest.fixture(scope="module")
def c0():
    wl = W.diatom_atom()
    orc = O.oracle_for_workload(wl)
    M, D, S = orc.matrix_make(wl.basis.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    b = wl.basis
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, M, D, wl.mass_amu * AMUTOAU)
    return wl, M, D, dyn


def compare(q_gpu, q_ora, N, amp_tol, cl_tol):
    a = np.max(np.abs(q_gpu[:2 * N] - q_ora[:2 * N]))
    c = np.max(np.abs(q_gpu[2 * N:2 * N + 6] - q_ora[2 * N:2 * N + 6]) /
               np.maximum(np.abs(q_ora[2 * N:2 * N + 6]), 1e-6))
    assert a <= amp_tol, f"amplitude diff {a}"
    assert c <= cl_tol, f"classical diff {c}"
    return a, c


@pytest.mark.parametrize("mode", [L.MQCT_PROP_CC_RK4, L.MQCT_PROP_CC_ODEINT])
def test_cc_trajectories_match_oracle(dev, c0, mode):
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    assert len(ls) == 83  # L = 0..82 at U=500 cm^-1, mu=3.6 amu, b<=15
    pick = [0, 23, 47, 71, 82]
    rows = prop.trajectories([ls[i] for i in pick], U_CM, s_st=1, time_step=10.0, eps=1e-8, hmin=1e-4)
    q, st = prop.run(mode, rows)
    N = b.n_states
    tol = (1e-9, 1e-9) if mode == L.MQCT_PROP_CC_RK4 else (1e-6, 1e-6)
    for i, r in enumerate(rows):
        qo, so, _ = dyn.dotraject(mode, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"], hmin=r["hmin"],
                                  R_start=r["R_start"])
        assert st["fail"][i] == so["fail"] == 0
        assert st["n_steps"][i] == so["n_steps"] or mode == L.MQCT_PROP_CC_ODEINT
        compare(q[i], qo, N, *tol)
        assert st["norm_err"][i] < 1e-6


@pytest.mark.parametrize("mode", [L.MQCT_PROP_AT_RK4, L.MQCT_PROP_AT_ODEINT])
def test_at_trajectories_match_oracle(dev, c0, mode):
    """AT-MQCT on paths recorded by the oracle's CC run (the first AT step)."""
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    pick = [5, 40, 80]
    paths = []
    for i in pick:
        r = prop.trajectories([ls[i]], U_CM, s_st=1, time_step=10.0)[0]
        _, _, rec = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, r["l_real"], r["E_coll"], r["dt"],
                                  R_start=r["R_start"], rec_cap=20000)
        paths.append(rec)
    prop.load_paths(paths)
    rows = prop.trajectories([ls[i] for i in pick], U_CM, s_st=1, time_step=10.0, eps=1e-8, hmin=1e-4,
                             at_paths=list(range(len(pick))))
    q, st = prop.run(mode, rows)
    N = b.n_states
    tol = 1e-9 if mode == L.MQCT_PROP_AT_RK4 else 1e-6
    for i, r in enumerate(rows):
        dyn.set_path(paths[i])
        qo, so, _ = dyn.dotraject(mode, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"], hmin=r["hmin"],
                                  R_start=r["R_start"])
        a = np.max(np.abs(q[i, :2 * N] - qo[:2 * N]))
        assert a <= tol, f"amplitude diff {a}"


def test_recorded_paths_match_oracle(dev, c0):
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    rows = prop.trajectories([ls[30]], U_CM, s_st=1, time_step=10.0)
    q, st, recs = prop.run_record(rows, 20000)
    _, _, rec_o = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, rows[0]["l_real"], rows[0]["E_coll"], rows[0]["dt"],
                                R_start=rows[0]["R_start"], rec_cap=20000)
    assert recs[0].shape == rec_o.shape
    assert np.max(np.abs(recs[0][:, :7] - rec_o[:, :7]) / np.maximum(np.abs(rec_o[:, :7]), 1e-6)) < 1e-9
    assert np.max(np.abs(recs[0][:, 7] - rec_o[:, 7])) < 1e-9 * np.abs(rec_o[:, 7]).max()
