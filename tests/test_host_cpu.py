"""CPU-side checks of the product: the host mirror's channel lists, state
tables and pair lists against the oracle (bit-exact), its eigenvectors
(a different solver) against the oracle's, the C-ABI library's export list
against include/mqct_hip.h, and the MTRX_UF.DAT writer/reader against an
independent byte-level parse of the Fortran record layout."""
import os
import re
import struct
import tempfile

import numpy as np
import pytest

import oracle.oracle as O
from mqct_amd import basis as B
from mqct_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_tables_match_oracle_type1():
    b = B.ini_arrays(B.channels_type1(0, 10, 1.93))
    ch = O.channels_type1(0, 10, 1.93)
    assert np.array_equal(ch["j"], b.j1_ch) and np.array_equal(ch["E"], b.E_ch)
    o = O.ini_arrays(1, b.j1_ch)
    for key in ("j12", "m12", "indx_chann", "chann_indx", "ind_mat"):
        assert np.array_equal(o[key], getattr(b, key)), key


def test_host_tables_match_oracle_type4():
    b = B.ini_arrays(B.channels_type4(0, 10, 27.9, 14.5, 9.3))
    ch = O.channels_type4(0, 10, 27.9, 14.5, 9.3)
    assert np.array_equal(ch["j"], b.j1_ch)
    assert np.array_equal(ch["ka"], b.ka1_ch) and np.array_equal(ch["kc"], b.kc1_ch)
    assert np.max(np.abs(ch["E"] - b.E_ch)) < 1e-10
    o = O.ini_arrays(4, b.j1_ch)
    for key in ("j12", "m12", "indx_chann", "chann_indx", "ind_mat"):
        assert np.array_equal(o[key], getattr(b, key)), key
    for c in range(b.n_channels):
        j = int(b.j1_ch[c])
        _, z = O.eigen_vec(27.9, 14.5, 9.3, j, int(b.ka1_ch[c]), int(b.kc1_ch[c]))
        assert np.max(np.abs(z - b.eig1[c, :2 * j + 1])) < 1e-12


def test_host_tables_match_oracle_type0():
    b = B.ini_arrays(B.channels_type0(0, 2, 0, 2, 27.9, 14.5, 9.3, 27.9, 14.5, 9.3))
    ch = O.channels_type0(0, 2, 0, 2, 27.9, 14.5, 9.3, 27.9, 14.5, 9.3)
    for a, c in (("j1", "j1_ch"), ("ka1", "ka1_ch"), ("kc1", "kc1_ch"), ("j2", "j2_ch"), ("ka2", "ka2_ch"),
                 ("kc2", "kc2_ch")):
        assert np.array_equal(ch[a], getattr(b, c))
    o = O.ini_arrays(0, b.j1_ch, b.j2_ch)
    for key in ("j12", "m12", "indx_chann", "chann_indx", "ind_mat"):
        assert np.array_equal(o[key], getattr(b, key)), key


def header_functions():
    txt = open(os.path.join(ROOT, "include", "mqct_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mqct_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import ctypes
    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libmqct_hip.so is not built (python -m mqct_amd.build)")
    lib = ctypes.CDLL(L.LIB_PATH)
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def parse_records(path):
    data = open(path, "rb").read()
    recs, p = [], 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        body = data[p + 4:p + 4 + n]
        (n2,) = struct.unpack_from("<i", data, p + 4 + n)
        assert n == n2
        recs.append(body)
        p += 8 + n
    return recs


def test_mtrx_uf_layout_and_round_trip():
    from mqct_amd import workloads as W
    from mqct_amd.matrix import read_mtrx_uf, write_mtrx_uf
    wl = W.asym_top_atom(jmax=2, n_r=6)
    b = wl.basis
    M = np.random.default_rng(5).standard_normal((b.total_size, len(wl.R)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "MTRX_UF.DAT")
        write_mtrx_uf(path, wl, M)
        recs = parse_records(path)
        # PRINT_MIJ_USER (matrix_ini.f:6658-6769): 5 header records, one per
        # state (st, chann, j12, m12, j, ka, kc), one per pair (k, st1, st2),
        # one per R point (i, R), then (k) and (Mat_el(:,k)) per element.
        hdr = [struct.unpack("<i", r)[0] for r in recs[:5]]
        assert hdr == [4, b.n_channels, b.n_states, b.total_size, len(wl.R)]
        p = 5
        for s in range(b.n_states):
            v = struct.unpack("<7i", recs[p + s])
            c = b.indx_chann[s]
            assert v == (s + 1, c, b.j12[s], b.m12[s], b.j1_ch[c - 1], b.ka1_ch[c - 1], b.kc1_ch[c - 1])
        p += b.n_states
        for k in range(b.total_size):
            assert struct.unpack("<3i", recs[p + k]) == (k + 1, *b.ind_mat[k])
        p += b.total_size
        for i in range(len(wl.R)):
            ii, r = struct.unpack("<id", recs[p + i])
            assert ii == i + 1 and r == wl.R[i]
        p += len(wl.R)
        for k in range(b.total_size):
            assert struct.unpack("<i", recs[p + 2 * k])[0] == k + 1
            assert np.array_equal(np.frombuffer(recs[p + 2 * k + 1], "<f8"), M[k])
        assert p + 2 * b.total_size == len(recs)
        back = read_mtrx_uf(path)
        assert np.array_equal(back["mat_el"], M) and np.array_equal(back["ind_mat"], b.ind_mat)
        assert np.array_equal(back["j12"], b.j12) and np.array_equal(back["R"], wl.R)
        assert np.array_equal(back["chan_qn"], b.chan_qn())


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: asked for more it exits with status 2 and
    says why, before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_cpu_baseline_sample_is_nearest_to_the_mean():
    """The CPU baseline times the trajectories whose right-hand-side counts
    are closest to the batch mean."""
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.baseline_sample(np.array([10, 50, 30, 31, 29, 100]), 4) == [1, 3, 2, 4]
    assert bench.baseline_sample(np.array([7]), 4) == [0]


def test_at_path_files_round_trip_reference_bytes(tmp_path):
    """The reference's own AT path files (EXAMPLES_IO/SYS_TYPE_0/
    H2O_H2O_AT_Approximation/AT_MQCT_2nd_Step/AT_APPROX_TRAJS/500.00000/0/0/
    AT_0_0_256_0 and n_AT_0_0_256_0, copied as data to tests/golden/at_files):
    the reader gets the count of n_AT and 8 REAL*8 per record (a path from
    R = 30 bohr inward and out, t increasing), and the writer reproduces
    both files byte for byte (propagator.f:2838-2852, 3457-3462, 4330-4333)."""
    from mqct_amd.propagate import read_at_path, write_at_paths
    src = os.path.join(os.path.dirname(__file__), "golden", "at_files")
    rec = read_at_path(src, 0, 0, 256, 0)
    assert rec.shape == (2227, 8)
    assert np.all(np.diff(rec[:, 0]) > 0)
    assert rec[0, 0] == 0.0 and rec[0, 1] == 30.0
    assert rec[:, 1].min() < 30.0
    write_at_paths(tmp_path, [rec], 0, 0, [256], 0)
    for nm in ("AT_0_0_256_0", "n_AT_0_0_256_0"):
        assert (tmp_path / nm).read_bytes() == open(os.path.join(src, nm), "rb").read(), nm
