"""GPU parity of the PES-expansion path (CALC_EXPANSION_TERMS and
EXPANSION_MATRIX_ELEMENT through the HIP library's C ABI) against the
oracle's restatement (oracle/mqct_oracle_exp.c, pinned in
test_expansion_oracle.py), on the same V grids and coefficients.

Bars (FP64 throughout): projections within 1e-12 of the largest
coefficient; matrix elements within 1e-12 of the largest element of the
build (the summation orders differ: MFMA tiles against the reference's
term loop); spline second derivatives within 1e-9 of the largest one (the
tridiagonal sweep amplifies the element differences); the skipped mask
identical.
"""
import numpy as np
import pytest

import oracle.oracle as O
from mqct_amd import _lib as L
from mqct_amd import expansion as X
from mqct_amd import workloads as W
from mqct_amd.matrix import MatrixBuilder

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind == 1:
        return W.diatom_atom(jmax=6, n_beta=32, n_r=24), X.terms_type1(8)
    if kind == 4:
        return W.asym_top_atom(jmax=3, n_beta=24, n_gamma=32, n_r=24), X.terms_type4(6)
    return W.two_asym_tops(2, 2, n_ang=8, n_r=3), X.terms_type0(2, 2, l_max=2)


@pytest.mark.parametrize("kind", [1, 4, 0])
def test_projection_matches_oracle(dev, kind):
    wl, terms = _case(kind)
    orc = O.oracle_for_workload(wl)
    V = np.array(orc.V, copy=True)
    MatrixBuilder(dev, wl, vgrid=V)
    ex = X.Expansion(dev, wl, terms)
    P = ex.project()
    Po = orc.expansion_terms(terms)
    d = np.max(np.abs(P - Po))
    s = np.max(np.abs(Po))
    print(f"type {kind}: {len(terms)} terms, projection diff {d:.2e} of {s:.2e}")
    assert s > 0 and d <= 1e-12 * s


@pytest.mark.parametrize("kind", [1, 4, 0])
def test_expansion_build_matches_oracle(dev, kind):
    wl, terms = _case(kind)
    if kind == 0:
        wl = W.two_asym_tops(2, 1, n_ang=4, n_r=40)
    rng = np.random.default_rng(11 + kind)
    # coefficients of a decaying potential with a sign change, cm^-1
    R = np.asarray(wl.R)
    C = rng.normal(size=(len(terms), len(R))) * (300.0 * np.exp(-0.9 * (R - 4.0)))[None, :]
    orc = O.oracle_for_workload(wl, np.zeros(len(R) * wl.grid.G))
    b = wl.basis
    ir1, ir2 = wl.ir_cut
    Mo, Do, So = orc.expansion_matrix_make(terms, C, b.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, ir1, ir2)
    MatrixBuilder(dev, wl)
    ex = X.Expansion(dev, wl, terms)
    M, D, S = ex.build(coeffs=C, R=R)
    dm = np.max(np.abs(M - Mo)) / np.max(np.abs(Mo))
    dd = np.max(np.abs(D - Do)) / np.max(np.abs(Do))
    print(f"type {kind}: {b.total_size} pairs, {len(terms)} terms, element diff {dm:.2e}, y'' diff {dd:.2e}, "
          f"skipped {int(So.sum())}")
    assert np.array_equal(S, So)
    assert dm <= 1e-12 and dd <= 1e-9


def test_projection_then_build_equals_given_coefficients(dev):
    wl, terms = _case(4)
    mb = MatrixBuilder(dev, wl)
    ex = X.Expansion(dev, wl, terms)
    P = ex.project()
    M1, D1, S1 = ex.build()
    M2, D2, S2 = ex.build(coeffs=P, R=np.asarray(wl.R))
    assert np.array_equal(M1, M2) and np.array_equal(D1, D2) and np.array_equal(S1, S2)
    t = ex.timing()
    assert all(v >= 0 for v in t.values())
    del mb


def test_gpu_expansion_equals_gpu_quadrature_for_potential_in_span(dev):
    """Type 0: V a finite sum of the terms' functions (from the oracle's
    term_grid): the HIP expansion build from the HIP projection equals the
    HIP quadrature build of the same V grid (1e-11 of the largest element;
    the grid integrates these products to ~1e-13)."""
    wl = W.two_asym_tops(1, 1, n_ang_each=(6, 16, 16, 6, 6), n_r=12)
    terms = [(0, 0, 0, 0, 0), (1, 0, 0, 0, 1), (1, 0, 1, 0, 2), (2, 0, 1, 0, 1), (1, 1, 1, 0, 0),
             (1, 1, 1, -1, 0), (1, 1, 1, 1, 2), (2, 2, 0, 0, 2), (2, 1, 1, 1, 1), (2, 2, 2, -2, 2)]
    o = O.oracle_for_workload(wl, np.zeros(len(wl.R) * wl.grid.G))
    F = np.array([o.term_grid(t) for t in terms])
    R = np.asarray(wl.R)
    c = np.random.default_rng(21).normal(size=(len(terms), len(R))) * (200.0 * np.exp(-0.7 * (R - 4.0)))[None, :]
    V = np.einsum("tr,tg->rg", c, F).ravel()
    mb = MatrixBuilder(dev, wl, vgrid=V)
    Mq, Dq, Sq = mb.build("fp64")
    ex = X.Expansion(dev, wl, terms)
    P = ex.project()
    assert np.max(np.abs(P - np.where([(t[1] == 0 and t[3] == 0) for t in terms], 1.0, 0.5)[:, None] * c)) \
        <= 1e-12 * np.max(np.abs(c))
    Me, De, Se = ex.build()
    d = np.max(np.abs(Me - Mq)) / np.max(np.abs(Mq))
    print(f"expansion vs quadrature on the GPU: {d:.2e}")
    assert d <= 1e-11


def test_bad_term_lists_are_rejected(dev):
    wl, _ = _case(0)
    MatrixBuilder(dev, wl)
    with pytest.raises(L.MqctError, match="M1 should be always positive"):
        X.Expansion(dev, wl, [(1, -1, 1, 0, 0)])
    with pytest.raises(L.MqctError, match="M2 should be positive only for M1 = 0"):
        X.Expansion(dev, wl, [(1, 0, 1, -1, 0)])
    with pytest.raises(L.MqctError, match="TRINAGULAR"):
        X.Expansion(dev, wl, [(1, 0, 1, 0, 3)])
    with pytest.raises(L.MqctError, match="Found two same terms"):
        X.Expansion(dev, wl, [(1, 1, 1, 1, 2), (1, 1, 1, 1, 2)])
    with pytest.raises(L.MqctError, match="Found two same terms"):
        X.Expansion(dev, wl, [(1, 1, 1, 1, 2), (1, -1, 1, -1, 2)])


def test_c3_basis_sampled_against_oracle(dev):
    """The benchmark's two-tops basis (2043 states, 282 179 pairs) with the
    312 terms of L1, L2 <= 4 and even NJU (a C2 axis on both partners):
    a sample of 600 elements against the oracle's EXPANSION_MATRIX_ELEMENT
    from the same coefficients (1e-12 of the sample's largest element)."""
    wl = W.two_asym_tops(W.C3_JMAX, W.C3_JMAX, emax=W.C3_EMAX, n_ang=2, n_r=24)
    terms = X.terms_type0(4, 4, du1=2, du2=2)
    assert len(terms) == 312 and wl.basis.n_states == 2043
    R = np.asarray(wl.R)
    C = np.random.default_rng(31).normal(size=(len(terms), len(R))) * (300.0 * np.exp(-0.9 * (R - 4.0)))[None, :]
    MatrixBuilder(dev, wl)
    ex = X.Expansion(dev, wl, terms)
    M, D, S = ex.build(coeffs=C, R=R)
    orc = O.oracle_for_workload(wl, np.zeros(len(R) * wl.grid.G))
    b = wl.basis
    ks = np.random.default_rng(32).choice(b.total_size, 600, replace=False)
    worst = scale = 0.0
    for k in ks:
        mo = orc.expansion_element(terms, C, b.ind_mat, int(k)) * wl.conv_unit_e
        if S[k]:
            assert max(abs(mo[wl.ir_cut[0] - 1]), abs(mo[wl.ir_cut[1] - 1])) < wl.mij_zero_cut
            continue
        worst = max(worst, np.max(np.abs(M[k] - mo)))
        scale = max(scale, np.max(np.abs(mo)))
    t = ex.timing()
    print(f"c3 basis: {b.total_size} pairs x {len(terms)} terms: angular {t['angular_ms']:.1f} ms, "
          f"GEMM {t['gemm_ms']:.1f} ms; sampled diff {worst:.2e} of {scale:.2e}")
    assert scale > 0 and worst <= 1e-12 * scale


def _golden_identical_workload(signs):
    import closed_forms as CF
    from mqct_amd import basis as B
    g = CF.h2o_h2o_identical_reference_system()
    b = B.channels_type0_list(g["rows"], g["A"], g["B"], g["C"])
    for i, r in enumerate(g["rows"]):
        b.eig1[i] *= signs[r[:3]]
        b.eig2[i] *= signs[r[3:]]
    B.ini_arrays(b, identical=True)
    base = W.two_asym_tops(1, 1, n_ang=2, n_r=2)
    wl = W.Workload("h2o_h2o_identical", b, base.grid, np.asarray(g["R"]), base.pes_kind, base.pes_par, 9.00764,
                    conv_unit_e=g["conv_unit_e"])
    return g, wl


def test_identical_h2o_h2o_matches_reference_mtrx_dat(dev):
    """The reference's IDENTICAL=YES H2O+H2O example through the HIP
    expansion build: all 644 Mij of its MTRX.DAT within 1e-9 of the largest
    (eigenvector signs as the CPU pin fits them), the oracle's values within
    1e-13 of the largest, and the cutoff mask equal to the file's zero
    elements."""
    signs = {(0, 0, 0): 1.0, (1, 1, 1): -1.0, (2, 0, 2): 1.0}
    g, wl = _golden_identical_workload(signs)
    MatrixBuilder(dev, wl)
    ex = X.Expansion(dev, wl, g["terms"])
    M, D, S = ex.build(coeffs=g["coeffs"], R=g["R"])
    scale = np.max(np.abs(g["mij"]))
    d = np.max(np.abs(M - g["mij"]))
    b = wl.basis
    grid = {"n_beta": 1, "n_gamma": 1, "n_alpha2": 1, "n_beta2": 1, "n_gamma2": 1}
    o = O.MatrixOracle(0, b.j12, b.m12, b.indx_chann, b.j1_ch, grid, g["R"], np.zeros(1), eig1=b.eig1, eig2=b.eig2,
                       j2_ch=b.j2_ch)
    o.set_identical(b)
    Mo = np.array([o.expansion_element(g["terms"], g["coeffs"], b.ind_mat, k) for k in range(b.total_size)]) \
        * g["conv_unit_e"]
    do = np.max(np.abs(M - Mo))
    print(f"identical H2O+H2O on the GPU: vs MTRX.DAT {d:.2e}, vs oracle {do:.2e} of {scale:.2e}")
    # the file's 26 elements that vanish by symmetry are the ones the cutoff skips
    assert np.array_equal(S.astype(bool), (g["mij"] == 0.0).all(axis=1)) and S.sum() == 26
    assert d <= 1e-9 * scale and do <= 1e-13 * scale


def test_identical_basis_refused_by_quadrature_build(dev):
    signs = {(0, 0, 0): 1.0, (1, 1, 1): 1.0, (2, 0, 2): 1.0}
    _, wl = _golden_identical_workload(signs)
    mb = MatrixBuilder(dev, wl)
    with pytest.raises(L.MqctError, match="identical partners"):
        mb.build("fp64")


def test_identical_pes_terms_match_oracle(dev):
    """IDENTICAL_PES=YES with distinguishable partners: the second symmetry
    term (the term with the partners' expansions exchanged) on the GPU
    against the oracle, 1e-12 of the largest element."""
    wl = W.two_asym_tops(2, 1, n_ang=2, n_r=24)
    terms = [t for t in X.terms_type0(2, 2, l_max=2) if t[0] > t[2] or (t[0] == t[2] and abs(t[1]) >= abs(t[3]))]
    R = np.asarray(wl.R)
    C = np.random.default_rng(41).normal(size=(len(terms), len(R))) * (300.0 * np.exp(-0.9 * (R - 4.0)))[None, :]
    MatrixBuilder(dev, wl)
    ex = X.Expansion(dev, wl, terms, identical_pes=True)
    M, D, S = ex.build(coeffs=C, R=R)
    o = O.oracle_for_workload(wl, np.zeros(len(R) * wl.grid.G))
    o.set_identical(wl.basis, identical=False, identical_pes=True)
    b = wl.basis
    Mo, Do, So = o.expansion_matrix_make(terms, C, b.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    o.set_identical(wl.basis, identical=False, identical_pes=False)
    M1 = np.array([o.expansion_element(terms, C, b.ind_mat, k) for k in range(0, b.total_size, 97)]) * wl.conv_unit_e
    d = np.max(np.abs(M - Mo)) / np.max(np.abs(Mo))
    print(f"IDENTICAL_PES: {len(terms)} terms, element diff {d:.2e}")
    assert d <= 1e-12 and np.array_equal(S, So)
    # the second symmetry term contributes
    assert np.max(np.abs(M1 - Mo[::97])) > 1e-3 * np.max(np.abs(Mo))
