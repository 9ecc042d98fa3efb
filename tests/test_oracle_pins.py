"""Pins of the CPU oracle: every function checked against something the
oracle does not compute itself -- closed forms, textbook routines,
mathematical invariants, exact reductions between system types, and the
reference's own example outputs (tests/golden/fixtures.json, extracted by
tests/golden/make_fixtures.py from EXAMPLES_IO).
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy import special

import oracle.oracle as O
from mqct_amd import basis as B
from mqct_amd import workloads as W

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


# -- special functions --------------------------------------------------------
def test_gauleg_matches_textbook():
    for n in (1, 2, 5, 24, 64, 101):
        x, w = O.gauleg(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) < 1e-14
        assert np.max(np.abs(w - wr)) < 1e-13
        assert abs(w.sum() - 2.0) < 1e-13


def test_plgndr_matches_scipy():
    rng = np.random.default_rng(0)
    for _ in range(200):
        l = int(rng.integers(0, 15))
        m = int(rng.integers(0, l + 1))
        x = rng.uniform(-1, 1)
        ref = special.lpmv(m, l, x)  # includes the Condon-Shortley phase
        assert abs(O.plgndr(l, m, x) - ref) <= 1e-11 * max(1.0, abs(ref))


def exact_3j(j1, j2, j3, m1, m2, m3):
    """Racah formula in exact rationals (integer j)."""
    if m1 + m2 + m3 != 0 or j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    pre = tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    s = Fraction(0)
    for k in range(0, j1 + j2 + j3 + 1):
        d = [k, j1 + j2 - j3 - k, j1 - m1 - k, j2 + m2 - k, j3 - j2 + m1 + k, j3 - j1 - m2 + k]
        if min(d) < 0:
            continue
        den = 1
        for x in d:
            den *= f(x)
        s += Fraction((-1) ** k, den)
    sign = (-1) ** (j1 - j2 - m3)
    return sign * math.sqrt(pre) * float(s)


def test_w3js_and_cg_closed_form():
    rng = np.random.default_rng(1)
    n = 0
    while n < 300:
        j1, j2 = int(rng.integers(0, 7)), int(rng.integers(0, 7))
        j3 = int(rng.integers(abs(j1 - j2), j1 + j2 + 1))
        m1, m2 = int(rng.integers(-j1, j1 + 1)), int(rng.integers(-j2, j2 + 1))
        m3 = -m1 - m2
        if abs(m3) > j3:
            continue
        ref = exact_3j(j1, j2, j3, m1, m2, m3)
        assert abs(O.w3js(2 * j1, 2 * j2, 2 * j3, 2 * m1, 2 * m2, 2 * m3) - ref) < 1e-13
        # CG <j1 m1 j2 m2|j3 M> = (-1)^(j1-j2+M) sqrt(2j3+1) (j1 j2 j3; m1 m2 -M)
        M = m1 + m2
        cg_ref = (-1) ** (j1 - j2 + M) * math.sqrt(2 * j3 + 1) * exact_3j(j1, j2, j3, m1, m2, -M)
        assert abs(O.cg(j1, j2, j3, m1, m2, M) - cg_ref) < 1e-13
        n += 1


def wigner_d_explicit(j, mp, m, b):
    f = math.factorial
    s = 0.0
    for k in range(0, 2 * j + 1):
        a = [j + m - k, k, mp - m + k, j - mp - k]
        if min(a) < 0:
            continue
        s += ((-1) ** (mp - m + k) * math.sqrt(f(j + mp) * f(j - mp) * f(j + m) * f(j - m)) /
              (f(a[0]) * f(a[1]) * f(a[2]) * f(a[3])) * math.cos(b / 2) ** (2 * j + m - mp - 2 * k) *
              math.sin(b / 2) ** (mp - m + 2 * k))
    return s


def test_djkm_is_wigner_d_of_m_k():
    """Djkm(j,k,m,alpha,beta,gamma) = d^j_{m k}(beta) exp(i(k gamma + m alpha))
    (explicit Wigner sum formula), all j<=6."""
    for b in (0.1, 0.7, 1.9, 3.0):
        for j in range(0, 7):
            for k in range(-j, j + 1):
                for m in range(-j, j + 1):
                    re, im = O.djkm(j, k, m, 0.4, b, 1.3)
                    d = wigner_d_explicit(j, m, k, b)
                    ph = k * 1.3 + m * 0.4
                    assert abs(re - d * math.cos(ph)) < 1e-12
                    assert abs(im - d * math.sin(ph)) < 1e-12


def test_spherical_harmonics_orthonormal_on_gauss_grid():
    x, w = O.gauleg(32)
    th = np.arccos(-x[::-1]) if False else (np.pi / 2 + np.pi / 2 * x)
    for m in (-3, 0, 2):
        rows = []
        for j in range(abs(m), 9):
            rows.append([O.wf_diat_top(j, m, t, 0.0)[0] for t in th])
        Y = np.array(rows)
        G = (Y * (w * np.sin(th) * np.pi / 2 * 2 * np.pi)) @ Y.T
        assert np.max(np.abs(G - np.eye(len(rows)))) < 1e-13


# -- asymmetric top against the reference's example data ----------------------
def test_rotor_levels_match_reference_examples():
    for key in ("h2o_he", "mf_he"):
        f = FIX[key]
        for j, ka, kc, e in f["channels"]:
            E, _ = O.eigen_vec(f["A"], f["B"], f["C"], j, ka, kc)
            assert abs(E - e) < 6e-4, (key, j, ka, kc, E, e)


def test_rotor_eigenvectors_match_user_basis_up_to_sign():
    f = FIX["h2o_he"]
    for (j, ka, kc, e), vec in zip(f["channels"], f["eigvecs"]):
        _, z = O.eigen_vec(f["A"], f["B"], f["C"], j, ka, kc)
        v = np.array(vec)
        s = np.sign(v[np.argmax(np.abs(v))]) * np.sign(z[np.argmax(np.abs(z))])
        assert np.max(np.abs(s * z - v)) < 1e-11


def test_rotor_eigen_invariants():
    for J in range(0, 11):
        w, z = O.asym_top_eigen(27.9, 14.5, 9.3, J)
        H = B.rotor_matrix(27.9, 14.5, 9.3, J)
        assert np.max(np.abs(H @ z - z * w)) < 1e-10
        assert np.max(np.abs(z.T @ z - np.eye(2 * J + 1))) < 1e-12
        assert np.all(np.diff(w) >= 0)


# -- state tables and pair lists: bit-exact against MTRX.DAT ------------------
@pytest.mark.parametrize("key", ["h2o_he", "mf_he"])
def test_state_tables_match_mtrx_dat(key):
    f = FIX[key]
    j_ch = np.array([c[0] for c in f["channels"]], np.int32)
    o = O.ini_arrays(4, j_ch)
    st = np.array(f["states"])
    assert len(o["j12"]) == f["header"]["NUMBER_OF_STATES"]
    assert np.array_equal(st[:, 0], np.arange(1, len(st) + 1))
    assert np.array_equal(st[:, 1], o["indx_chann"])
    assert np.array_equal(st[:, 2], o["j12"])
    assert np.array_equal(st[:, 3], o["m12"])
    assert np.array_equal(st[:, 4], j_ch[o["indx_chann"] - 1])
    pairs = np.array(f["pairs"])
    assert len(o["ind_mat"]) == f["header"]["MATRIX_SIZE"]
    assert np.array_equal(pairs[:, 1:], o["ind_mat"])


def test_state_counts_of_the_workloads():
    """Closed-form counts: type 1 j<=10: sum(2j+1)=121 states; type 4
    j<=10: sum (2j+1)^2 = 1771; type 0 j1,j2<=2: (sum_j (2j+1)^2)^2 = 1225."""
    o = O.ini_arrays(1, np.arange(11))
    assert len(o["j12"]) == 121 and len(o["ind_mat"]) == sum(
        sum(1 for j in range(11) if j >= abs(m)) * (sum(1 for j in range(11) if j >= abs(m)) + 1) // 2
        for m in range(-10, 11))
    ch = O.channels_type4(0, 10, 27.9, 14.5, 9.3)
    o4 = O.ini_arrays(4, ch["j"])
    assert len(o4["j12"]) == 1771 and len(o4["ind_mat"]) == 86779
    ch0 = O.channels_type0(0, 2, 0, 2, 27.9, 14.5, 9.3, 27.9, 14.5, 9.3)
    o0 = O.ini_arrays(0, ch0["j1"], ch0["j2"])
    assert len(ch0["j1"]) == 81 and len(o0["j12"]) == 1225


# -- matrix elements ----------------------------------------------------------
def test_type1_matrix_is_gaunt_closed_form():
    """V = sum_l V_l(R) P_l(cos theta) (the core in V_0): <j m|V|j' m> = sum_l V_l (-1)^m
    sqrt((2j+1)(2j'+1)) (j l j';0 0 0)(j l j';-m 0 m); the 64-point Gauss
    grid is exact for these degrees."""
    wl = W.diatom_atom(n_r=5)
    b = wl.basis
    orc = O.oracle_for_workload(wl)
    par = wl.pes_par
    Lp = int(par[0])
    A, bb, C = par[1:1 + Lp], par[1 + Lp:1 + 2 * Lp], par[1 + 2 * Lp:1 + 3 * Lp]
    Aw, bw, Rw = par[1 + 3 * Lp:]
    for ir, R in enumerate(wl.R):
        Vl = A * np.exp(-bb * R) - C / R ** 6
        Vl[0] += Aw * np.exp(-bw * (R - Rw))  # isotropic core
        for k in range(0, b.total_size, 7):
            s1, s2 = b.ind_mat[k] - 1
            j1, j2, m = int(b.j12[s1]), int(b.j12[s2]), int(b.m12[s1])
            ref = sum(Vl[l] * (-1) ** m * math.sqrt((2 * j1 + 1) * (2 * j2 + 1)) * exact_3j(j1, l, j2, 0, 0, 0) *
                      exact_3j(j1, l, j2, -m, 0, m) for l in range(Lp))
            got = orc.element(b.ind_mat, k, ir)
            scale = np.abs(Vl).sum()
            assert abs(got - ref) <= 1e-12 * scale


def test_type4_reduces_to_type1():
    """Asymmetric-top states that are pure K=0 (c_K = delta_K0) in a
    gamma-independent potential give exactly the diatom-atom elements: the
    type-4 integrand and weights (pot_comp.f:2535-2557) against type 1
    (pot_comp.f:1947-1952)."""
    wl1 = W.diatom_atom(jmax=4, n_beta=20, n_r=3)
    b1 = wl1.basis
    nb, ng = 20, 8
    g4 = B.grid_type4(nb, ng)
    ld = 2 * 4 + 1
    eig = np.zeros((b1.n_channels, ld))
    for c, j in enumerate(b1.j1_ch):
        eig[c, j] = 1.0  # K = 0 slot
    o1 = O.oracle_for_workload(wl1)
    V1 = o1.V.reshape(3, nb)
    V4 = np.repeat(V1[:, :, None], ng, axis=2).ravel()
    o4 = O.MatrixOracle(4, b1.j12, b1.m12, b1.indx_chann, b1.j1_ch, O.grid_dict(g4), wl1.R, V4, eig1=eig)
    for k in range(b1.total_size):
        for ir in range(3):
            a, c = o1.element(b1.ind_mat, k, ir), o4.element(b1.ind_mat, k, ir)
            assert abs(a - c) <= 1e-13 * max(1.0, abs(a)) + 1e-12 * np.abs(V1[ir]).max()


def test_type0_reduces_to_type4():
    """Second top in its J=0 level and a potential independent of its
    angles: the pair functions (CG sum, basis_wave_functions.f:1-41) and the
    5-D weights (pot_comp.f:3718-3722) give exactly the type-4 elements."""
    A, Bc, Cc = 27.9, 14.5, 9.3
    b0 = B.ini_arrays(B.channels_type0(0, 2, 0, 0, A, Bc, Cc, A, Bc, Cc))
    b4 = B.ini_arrays(B.channels_type4(0, 2, A, Bc, Cc))
    assert np.array_equal(b0.j12, b4.j12) and np.array_equal(b0.ind_mat, b4.ind_mat)
    # 14 Gauss points integrate the J=0 Jacobian sin(beta2) to 1e-16
    g0 = B.grid_type0(3, 8, 14, 10, 4)
    g4 = B.grid_type4(8, 10)
    R = np.array([4.0, 6.0])
    rng = np.random.default_rng(3)
    V4 = rng.standard_normal((2, 8, 10))
    V0 = np.broadcast_to(V4[:, None, :, None, :, None], (2, 3, 8, 14, 10, 4)).ravel()
    o4 = O.MatrixOracle(4, b4.j12, b4.m12, b4.indx_chann, b4.j1_ch, O.grid_dict(g4), R, V4.ravel(),
                        eig1=b4.eig1)
    o0 = O.MatrixOracle(0, b0.j12, b0.m12, b0.indx_chann, b0.j1_ch, O.grid_dict(g0), R, V0,
                        eig1=b0.eig1, eig2=b0.eig2, j2_ch=b0.j2_ch)
    for k in range(b0.total_size):
        for ir in range(2):
            a, c = o4.element(b4.ind_mat, k, ir), o0.element(b0.ind_mat, k, ir)
            assert abs(a - c) <= 1e-12 * max(1.0, abs(a))


def test_spline_exact_on_cubics():
    x = np.sort(np.random.default_rng(4).uniform(0, 5, 17))
    x[0], x[-1] = 0.0, 5.0
    y = 0.3 * x ** 3 - x ** 2 + 2 * x - 1
    yp1 = 0.9 * x[0] ** 2 - 2 * x[0] + 2
    ypn = 0.9 * x[-1] ** 2 - 2 * x[-1] + 2
    y2 = O.spline(x, y, yp1, ypn)
    assert np.max(np.abs(y2 - (1.8 * x - 2))) < 1e-10
    for t in np.linspace(0.01, 4.99, 37):
        v, d = O.splint(x, y, y2, t)
        assert abs(v - (0.3 * t ** 3 - t ** 2 + 2 * t - 1)) < 1e-10
        assert abs(d - (0.9 * t ** 2 - 2 * t + 2)) < 1e-9


def test_cutoff_and_spline_of_matrix_make():
    """MATRIX_MAKE post-processing on the oracle's own elements: columns
    below MIJ_ZERO_CUT at both cutoff points are zeroed; the others carry
    the clamped spline with finite-difference end slopes."""
    wl = W.diatom_atom(n_r=40)
    b = wl.basis
    orc = O.oracle_for_workload(wl)
    M, D, S = orc.matrix_make(b.ind_mat, wl.conv_unit_e, 1e-4, *wl.ir_cut)
    ir1, ir2 = wl.ir_cut
    for k in range(b.total_size):
        raw = np.array([orc.element(b.ind_mat, k, i) * wl.conv_unit_e for i in range(40)])
        cut = max(abs(raw[ir1 - 1]), abs(raw[ir2 - 1])) < 1e-4
        assert S[k] == cut
        if cut:
            assert not M[k].any() and not D[k].any()
        else:
            assert np.array_equal(M[k], raw)
            y2 = O.spline(wl.R, raw, (raw[1] - raw[0]) / (wl.R[1] - wl.R[0]),
                          (raw[-1] - raw[-2]) / (wl.R[-1] - wl.R[-2]))
            assert np.array_equal(D[k], y2)


def test_mij_shift_of_both_branches():
    """MIJ_SHIFT (default YES, iotest.f:3754).  Compute branch
    (matrix_ini.f:2226-2230): every row subtracts the UNSHIFTED last row,
    so the last R point becomes exactly 0 and the others keep their
    differences exactly.  Read branch (matrix_ini.f:6603-6621): the same
    shift, then the spline of the shifted values; a cubic spline of y - c
    has the second derivatives of y, so both branches carry the same y''
    to rounding and the interpolant moves by exactly the constant.  A shift
    by any other row (or a row shifted before the others read it) breaks
    the zero at R_max or the differences."""
    wl = W.diatom_atom(n_r=40)
    b = wl.basis
    orc = O.oracle_for_workload(wl)
    M, D, S = orc.matrix_make(b.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    Ms = O.mij_shift(M)
    assert np.all(Ms[:, -1] == 0.0)
    assert np.array_equal(Ms[:, :-1], M[:, :-1] - M[:, -1:])
    assert not np.array_equal(Ms, M)
    Mr, Dr = O.read_mij_post(wl.R, M, shift=True)
    assert np.array_equal(Mr, Ms)
    M0, D0 = O.read_mij_post(wl.R, M, shift=False)
    assert np.array_equal(M0, M)
    # without the shift the read branch re-splines to the compute branch's y''
    live = S == 0
    assert np.array_equal(D0[live], D[live])
    scale = np.abs(D0).max(axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(Dr - D0) / scale) < 1e-8
    for k in np.nonzero(live)[0][:12]:
        for x in (3.3, 5.05, 9.9, 17.2, wl.R[-1]):
            v0, d0 = O.splint(wl.R, M[k], D0[k], x)
            v1, d1 = O.splint(wl.R, Mr[k], Dr[k], x)
            big = max(abs(v0), abs(M[k, -1]), 1e-30)
            assert abs((v1 + M[k, -1]) - v0) <= 1e-12 * big
            assert abs(d1 - d0) <= 1e-8 * (abs(d0) + np.abs(M[k]).max())
        assert O.splint(wl.R, Mr[k], Dr[k], wl.R[-1])[0] == 0.0


@pytest.mark.parametrize("ref_xz", [1, 0])
def test_type4_elements_closed_form(ref_xz):
    """Type-4 elements of K-mixed asymmetric-top states (j <= 3, the
    benchmark's rotor, eigenvectors with K != 0 components) for a potential
    built from D^l_{0q} (l <= 4, random complex weights, two R points):
    the oracle's INTEGRATOR_TEMP (pot_comp.f:2503-2557) against the 3j
    closed form of tests/closed_forms.py, in the XZ frame (real part) and
    with REF_XZ=NO (real + imaginary part, pot_comp.f:2544-2553).  The
    12 x 24 grid integrates these degrees exactly.  A sign slip in the
    REF_XZ=NO integrand, a transposed pair (psi_1^* psi_2) or a wrong K
    phase fails at the first element with Im != 0."""
    import closed_forms as CF
    A, Bc, Cc = 27.9, 14.5, 9.3
    b = B.ini_arrays(B.channels_type4(0, 3, A, Bc, Cc))
    nb, ng = 28, 24
    g = B.grid_type4(nb, ng, ref_xz=ref_xz)
    beta = g.xbm + g.xbl * np.asarray(g.xb)
    gamma = np.asarray(g.xg)
    rng = np.random.default_rng(41)
    ws = [CF.d_potential_coeffs(rng, 4) for _ in range(2)]
    V = np.stack([CF.d_potential_grid(w, beta, gamma) for w in ws])  # [R][beta][gamma]
    R = np.array([5.0, 7.0])
    o = O.MatrixOracle(4, b.j12, b.m12, b.indx_chann, b.j1_ch, O.grid_dict(g), R, V.ravel(), eig1=b.eig1)
    n_mixed = n_imag = 0
    for k in range(b.total_size):
        s1, s2 = b.ind_mat[k] - 1
        c1, c2 = b.indx_chann[s1] - 1, b.indx_chann[s2] - 1
        J1, J2, M = int(b.j1_ch[c1]), int(b.j1_ch[c2]), int(b.m12[s1])
        e1, e2 = b.eig1[c1, :2 * J1 + 1], b.eig1[c2, :2 * J2 + 1]
        n_mixed += int(np.count_nonzero(np.abs(e1) > 1e-12) > 1)
        for ir in range(2):
            ref = CF.type4_element(ws[ir], J1, e1, J2, e2, M, ref_xz=ref_xz)
            if ref_xz == 0:
                n_imag += int(abs(ref - CF.type4_element(ws[ir], J1, e1, J2, e2, M, ref_xz=1)) > 1e-6)
            got = o.element(b.ind_mat, k, ir)
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (k, ir, got, ref)
    assert n_mixed > 100
    if ref_xz == 0:
        assert n_imag > 100  # the imaginary part is present and pinned


def test_type4_matches_reference_mtrx_dat():
    """The reference's own SYS_TYPE 4 run (EXAMPLES_IO/SYS_TYPE_4/
    H2O_He_external_rot_wfn/MTRX.DAT: 147 elements x 2 R, direct quadrature
    with K-mixed eigenvectors given in USER_DEFINED_BASIS.DAT) reproduced by
    the oracle from the example's own PES on the example's grid: every
    element within 1e-12 of the largest (the file prints 12 digits), the
    significant ones within 1e-11 relative, and the MIJ_CUTOFF mask equal to
    the file's zero elements."""
    import closed_forms as CF
    b, g, R, V, ref = CF.h2ohe_reference_system()
    o = O.MatrixOracle(4, b.j12, b.m12, b.indx_chann, b.j1_ch, O.grid_dict(g), R, V.ravel(), eig1=b.eig1)
    M, D, S = o.matrix_make(b.ind_mat, B.CM_TO_HARTREE, 1e-12 * B.CM_TO_HARTREE, 1, 2)
    scale = np.abs(ref).max()
    assert np.array_equal(S.astype(bool), np.all(ref == 0.0, axis=1))
    assert np.max(np.abs(M - ref)) <= 1e-12 * scale
    big = np.abs(ref) > 1e-3 * scale
    assert big.sum() > 100
    assert np.max(np.abs(M[big] - ref[big]) / np.abs(ref[big])) <= 1e-11
