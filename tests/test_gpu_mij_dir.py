"""A split build's matrix directory through the GPU path: the HIP build
written as MATRIX_TRUNCATED, read back as bk_read_matrix reads it, then
MIJ_SHIFT and the spline on the GPU (mqct_hip_mij_spline): against the
oracle's build compacted to its non-zero elements and post-processed as
READ_MIJ_USER does (or_read_mij_post), 1e-10 of the largest element."""
import numpy as np
import pytest

import oracle.oracle as O
from mqct_amd import workloads as W
from mqct_amd.matrix import MatrixBuilder, mij_spline, read_mij_dir, write_mij_dir

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("unformatted", [False, True])
def test_build_write_read_spline(dev, tmp_path, unformatted):
    wl = W.asym_top_atom(jmax=4, n_beta=16, n_gamma=32, n_r=40)
    mb = MatrixBuilder(dev, wl)
    M, D, S = mb.build("fp64")
    out = str(tmp_path / "MATRIX_TRUNCATED")
    write_mij_dir(out, wl, M, n_files=9, unformatted=unformatted)
    ind, Mr, cut = read_mij_dir(out, len(wl.R), unformatted=unformatted)
    b = wl.basis
    assert np.array_equal(ind, b.ind_mat[S == 0])
    dev.set_mij_shift(True)
    Ms, Ds = mij_spline(dev, wl.R, Mr)
    orc = O.oracle_for_workload(wl)
    Mo, Do, So = orc.matrix_make(b.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    assert np.array_equal(S, So) and S.any() and not S.all()
    Mo, Do = O.read_mij_post(wl.R, Mo[So == 0], shift=True)
    s = np.abs(Mo).max()
    assert np.max(np.abs(Ms - Mo)) <= 1e-10 * s
    assert np.max(np.abs(Ds - Do)) <= 1e-8 * np.abs(Do).max()
