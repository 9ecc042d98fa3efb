"""MIJ_SHIFT on the GPU (the reference's default, iotest.f:3754) at both of
its sites, through the C ABI, against the oracle:

* mqct_hip_use_built_matrix: the compute branch (matrix_ini.f:2226-2230),
  values shifted after the spline.  CC-MQCT RK4 trajectories on the GPU's
  own FP64 build against the oracle's DOTRAJECT on the oracle's shifted
  matrix with the unshifted second derivatives (1e-9, as the unshifted
  trajectories); with the flag off, against the unshifted matrix.
* mqct_hip_mij_spline: READ_MIJ_USER's shift and re-spline
  (matrix_ini.f:6603-6621): values bit-exact (one subtraction), second
  derivatives within 1e-12 of the largest |y''| of the element.
"""
import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.basis import AMUTOAU, CM_TO_HARTREE
from mqct_amd.matrix import MatrixBuilder, mij_spline
from mqct_amd.propagate import Propagator
import oracle.oracle as O

pytestmark = pytest.mark.gpu

U_CM = 500.0


@pytest.fixture(scope="module")
def c0_oracle():
    wl = W.diatom_atom(n_r=80)
    orc = O.oracle_for_workload(wl)
    M, D, S = orc.matrix_make(wl.basis.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    return wl, M, D, S


@pytest.mark.parametrize("shift", [1, 0])
def test_read_branch_shift_and_spline(dev, c0_oracle, shift):
    wl, M, D, S = c0_oracle
    dev.set_mij_shift(shift)
    try:
        Mg, Dg = mij_spline(dev, wl.R, M)
    finally:
        dev.set_mij_shift(1)
    Mo, Do = O.read_mij_post(wl.R, M, shift=bool(shift))
    assert np.array_equal(Mg, Mo)
    if shift:
        assert np.all(Mg[:, -1] == 0.0)
    scale = np.abs(Do).max(axis=1, keepdims=True) + 1e-300
    err = np.max(np.abs(Dg - Do) / scale)
    print(f"shift {shift}: y'' max rel err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("shift", [1, 0])
def test_built_matrix_propagates_shifted_values(dev, c0_oracle, shift):
    wl, M, D, S = c0_oracle
    b = wl.basis
    mb = MatrixBuilder(dev, wl)
    Mg, Dg, Sg = mb.build("fp64")
    assert np.array_equal(Sg, S)
    # the built matrix itself is never shifted (the reference writes it first)
    assert np.max(np.abs(Mg - M) / np.maximum(np.abs(M), 1e-300)) <= 1e-10
    dev.set_mij_shift(shift)
    try:
        prop = Propagator(dev, b, wl.R, mass_amu=wl.mass_amu, from_build=True)
    finally:
        dev.set_mij_shift(1)
    ls, _, _ = prop.l_grid(U_CM, 15.0, 1)
    rows = prop.trajectories([ls[i] for i in (0, 30, 60, 82)], U_CM, s_st=1, time_step=10.0)
    q, st = prop.run(L.MQCT_PROP_CC_RK4, rows)
    Mo = O.mij_shift(M) if shift else M
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, Mo, D, wl.mass_amu * AMUTOAU)
    dyn_other = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, M if shift else O.mij_shift(M),
                            D, wl.mass_amu * AMUTOAU)
    N = b.n_states
    moved = 0.0
    for i, r in enumerate(rows):
        qo, so, _ = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, r["l_real"], r["E_coll"], r["dt"], R_start=r["R_start"])
        assert st["n_steps"][i] == so["n_steps"]
        d = np.max(np.abs(q[i, :2 * N] - qo[:2 * N]))
        assert d <= 1e-9, f"trajectory {i}: amplitude diff {d}"
        qx, _, _ = dyn_other.dotraject(L.MQCT_PROP_CC_RK4, 1, r["l_real"], r["E_coll"], r["dt"],
                                       R_start=r["R_start"])
        moved = max(moved, np.max(np.abs(qx[:2 * N] - qo[:2 * N])))
    # the shift changes the amplitudes far beyond the parity bar
    assert moved > 1e-6, moved
