"""The parallel-IO matrix directory (MATRIX_TRUNCATED) of a split build,
host code of the library (mqct_hip_write_mij_info / write_mij_dir /
read_mij_dir), pinned by the reference's own files:

  * EXAMPLES_IO/SYS_TYPE_6/CO_H2_Prob_spline/MATRIX_TRUNCATED (copied under
    tests/golden/mij_truncated): its 16 elements over 10 files are read,
    and written back by our writer with the same rank split, byte for byte
    (MIJ_*, Ind_*, MATRIX_NONZERO_INDEX.DAT, MATRIX_FILE_INDEX.DAT);
  * MATRIX_INFO.DAT against the same header lines of the reference's
    identical-partner H2O+H2O MTRX.DAT (states with the exchange-parity
    column, pairs, R grid), byte for byte;
  * formatted and unformatted round trips of a matrix with elements under
    the cut, over an uneven rank split.
"""
import os

import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.basis import CM_TO_HARTREE
from mqct_amd.matrix import read_mij_dir, write_mij_dir

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_reference_directory_reads_and_writes_back_byte_for_byte(tmp_path):
    src = os.path.join(GOLD, "mij_truncated")
    info = open(os.path.join(src, "MATRIX_INFO.DAT")).read().splitlines()
    n_r = int(info[4].split("=")[1])
    ind, M, cut = read_mij_dir(src, n_r)
    assert ind.shape == (16, 2) and M.shape == (16, n_r)
    assert cut == 1e-12
    assert M[0, 0] == 0.523442530681E-01 and M[12, -1] == -0.366862289188E-16
    # the pairs are those MATRIX_INFO.DAT lists (all 16 elements are above the cut)
    i0 = next(i for i, ln in enumerate(info) if "MIJ_INDEX" in ln)
    pairs = np.array([[int(x) for x in ln.split()[1:]] for ln in info[i0 + 1:i0 + 17]])
    assert np.array_equal(ind, pairs)
    R = np.array([float(ln.split()[1]) for ln in info[i0 + 18:i0 + 18 + n_r]])
    ir1, ir2 = int(np.argmin(np.abs(R - 6.0))) + 1, int(np.argmin(np.abs(R - 8.0))) + 1
    out = str(tmp_path)
    lib = L.load()
    L.check(lib.mqct_hip_write_mij_dir(out.encode(), 0, 10, 16, n_r, L.iptr(L.as_i32(ind)), L.dptr(L.as_f64(M)),
                                       1e-12 * CM_TO_HARTREE, ir1, ir2), "write_mij_dir")
    names = [f for f in sorted(os.listdir(src)) if f != "MATRIX_INFO.DAT"]
    assert sorted(f for f in os.listdir(out)) == names
    for f in names:
        a = open(os.path.join(src, f), "rb").read()
        b = open(os.path.join(out, f), "rb").read()
        assert a == b, f


def test_matrix_info_matches_reference_header(tmp_path):
    import closed_forms as CF
    from mqct_amd import basis as B
    g = CF.h2o_h2o_identical_reference_system()
    b = B.ini_arrays(B.channels_type0_list(g["rows"], g["A"], g["B"], g["C"]), identical=True)
    lib = L.load()
    out = str(tmp_path)
    R = L.as_f64(g["R"])
    L.check(lib.mqct_hip_write_mij_info(out.encode(), 0, 0, b.n_channels, b.n_states, b.total_size, len(R),
                                        L.iptr(L.as_i32(b.indx_chann)), L.iptr(L.as_i32(b.j12)),
                                        L.iptr(L.as_i32(b.m12)), L.iptr(b.chan_qn()),
                                        L.iptr(L.as_i32(b.parity_state)), L.iptr(L.as_i32(b.ind_mat)), L.dptr(R)),
            "write_mij_info")
    mine = open(os.path.join(out, "MATRIX_INFO.DAT")).read().splitlines()
    ref = open(os.path.join(GOLD, "expansion", "h2o_h2o_identical_MTRX.DAT")).read().splitlines()
    n = len(mine) - 1
    assert mine[:n] == ref[:n] and mine[-1] == " "
    assert ref[n].split() == ["#k_matel", "#iR", "Mij(iR,k_matel)"]


@pytest.mark.parametrize("unformatted", [0, 1])
def test_round_trip_with_cut_elements(tmp_path, unformatted):
    wl = W.asym_top_atom(jmax=2, n_beta=8, n_gamma=16, n_r=20)
    b = wl.basis
    rng = np.random.default_rng(7)
    M = rng.normal(size=(b.total_size, len(wl.R))) * 1e-4
    dead = rng.random(b.total_size) < 0.3
    M[dead] = 0.0
    M[np.nonzero(dead)[0][:3]] = 1e-20          # under the cut, not zero
    out = str(tmp_path / "MATRIX_TRUNCATED")
    write_mij_dir(out, wl, M, n_files=7, unformatted=bool(unformatted))
    ind, Mr, cut = read_mij_dir(out, len(wl.R), unformatted=bool(unformatted))
    keep = ~dead
    assert np.array_equal(ind, b.ind_mat[keep])
    if unformatted:
        assert np.array_equal(Mr, M[keep])
    else:
        assert np.max(np.abs(Mr - M[keep]) / np.abs(M[keep])) < 1e-11
    assert abs(cut - 1e-12) < 1e-24
    files = sorted(os.listdir(out))
    assert len([f for f in files if f.startswith("MIJ_")]) == 7
    with pytest.raises(L.MqctError, match="larger than #elements"):
        write_mij_dir(str(tmp_path / "x"), wl, M, n_files=b.total_size + 1)
