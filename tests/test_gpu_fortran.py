"""The Fortran host path: a flang-built program drives the library through
the iso_c_binding bridge (mqct_amd/fortran).  Checks: the MTRX_UF.DAT that
Fortran's own unformatted WRITEs produce in the PRINT_MIJ_USER layout is
byte-identical to the library's writer; the matrix the Fortran host got
matches the oracle to 1e-10 relative; the expansion path through the
bridge (projection and matrix) matches the oracle's to 1e-12; after the
driver's MIJ_SHIFT its CC-MQCT RK4 trajectories match the oracle's to
1e-9."""
import os
import subprocess

import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.basis import AMUTOAU, CM_TO_HARTREE
import oracle.oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "mqct_amd", "fortran", "_build", "mqct_fortran_demo")


def test_fortran_host_drives_the_library(tmp_path):
    if not os.path.exists(DEMO):
        pytest.skip("Fortran demo not built (no Fortran compiler when build() ran)")
    wl = W.diatom_atom(jmax=2, n_beta=16, n_r=12)
    par = wl.pes_par
    ls = [0.0, 12.0]
    nml = ["&demo", " jmax=2, n_beta=16, n_r=12, rmin=3.0, rmax=20.0, be=1.93,",
           " npar=%d," % len(par), " par=" + ", ".join("%.17g" % v for v in par) + ",",
           " e_coll_cm=500.0, mass_amu=3.6, time_step=20.0, n_traj=2,",
           " l_values=" + ", ".join("%.1f" % v for v in ls), "/"]
    (tmp_path / "demo.nml").write_text("\n".join(nml) + "\n")
    r = subprocess.run([DEMO, "demo.nml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f_bytes = (tmp_path / "MTRX_UF.DAT").read_bytes()
    c_bytes = (tmp_path / "MTRX_UF_C.DAT").read_bytes()
    assert f_bytes == c_bytes
    b = wl.basis
    T, n = b.total_size, len(wl.R)
    N = b.n_states
    raw = np.fromfile(tmp_path / "demo_out.bin", dtype=np.float64)
    nq = 2 * (2 * N + 8)
    assert raw.size == 4 * T * n + nq + 7 * n
    M = raw[:T * n].reshape(T, n)            # Mat_el(n_r, T) column-major
    D = raw[T * n:2 * T * n].reshape(T, n)
    q = raw[2 * T * n:2 * T * n + nq].reshape(2, 2 * N + 8)
    o0 = 2 * T * n + nq
    Me = raw[o0:o0 + T * n].reshape(T, n)
    De = raw[o0 + T * n:o0 + 2 * T * n].reshape(T, n)
    Pe = raw[o0 + 2 * T * n:].reshape(7, n)  # expansion_terms(n_r, 7)
    orc = O.oracle_for_workload(wl)
    Mo, Do, So = orc.matrix_make(b.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    # the demo applies the driver's MIJ_SHIFT (matrix_ini.f:2226-2230) before
    # propagating; the file it wrote before that holds the unshifted values
    Ms = O.mij_shift(Mo)
    assert np.all(M[:, -1] == 0.0)
    assert np.max(np.abs(M - Ms)) <= 1e-10 * np.abs(Mo).max()
    from mqct_amd.matrix import read_mtrx_uf
    Mf = read_mtrx_uf(str(tmp_path / "MTRX_UF.DAT"))["mat_el"]
    assert np.max(np.abs(Mf - Mo) / np.maximum(np.abs(Mo), 1e-300)) <= 1e-10
    assert np.max(np.abs(D - Do)) <= 1e-10 * np.abs(Do).max()
    # the expansion path through the bridge: projection and matrix against
    # the oracle's CALC_EXPANSION_TERMS and EXPANSION_MATRIX_ELEMENT
    terms = [(l,) for l in range(7)]
    Po = orc.expansion_terms(terms)
    assert np.max(np.abs(Pe - Po)) <= 1e-12 * np.abs(Po).max()
    Meo, Deo, _ = orc.expansion_matrix_make(terms, Po, b.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    assert np.max(np.abs(Me - Meo)) <= 1e-12 * np.abs(Meo).max()
    assert np.max(np.abs(De - Deo)) <= 1e-9 * np.abs(Deo).max()
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, M, D, 3.6 * AMUTOAU)
    dt = 20.0 / np.sqrt(500.0 / 300.0)
    for i, l in enumerate(ls):
        qo, so, _ = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, l, 500.0 * CM_TO_HARTREE, dt, eps=1e-8, hmin=1e-4,
                                  R_start=wl.R[-1])
        assert np.max(np.abs(q[i, :2 * N] - qo[:2 * N])) <= 1e-9
