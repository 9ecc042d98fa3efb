"""GPU parity of the matrix build (the HIP library through its C ABI) against
the CPU oracle, on seeded inputs.

Bar: FP64 elements within 1e-10 relative, except where the quadrature sum
cancels: an element whose terms sum in absolute value to A may differ by
1e-13*A (a few hundred roundings of the largest term -- any summation order
does that), so the test is |gpu - oracle| <= max(1e-10 |oracle|, 1e-13 A).
State tables, pair list and the cutoff mask are bit-exact.  Mixed-precision
forms are held to their worst-case rounding bound against A: 4e-6 A for
both FP32 and split-BF16x3; per-element errors are also reported relative
to the element and to its largest value over R.
"""
import numpy as np
import pytest

from mqct_amd import workloads as W
from mqct_amd.matrix import MatrixBuilder
import oracle.oracle as O

pytestmark = pytest.mark.gpu


def rel_err(g, o, floor_frac=1e-12):
    atol = floor_frac * np.abs(o).max()
    return np.max(np.abs(g - o) / np.maximum(np.abs(o), atol))


def fp64_ok(g, o, A):
    """max over elements of |g-o| / max(1e-10|o|, 1e-13 A); <= 1 passes."""
    den = np.maximum(np.maximum(1e-10 * np.abs(o), 1e-13 * A), 1e-300)
    return np.max(np.abs(g - o) / den)


_ABS_CACHE = {}


def abs_sums(orc, wl, ks=None, irs=None):
    whole = ks is None and irs is None
    if whole and _wl_key(wl) in _ABS_CACHE:
        return _ABS_CACHE[_wl_key(wl)]
    ks = range(wl.basis.total_size) if ks is None else ks
    irs = range(len(wl.R)) if irs is None else irs
    A = orc.abs_sums(wl.basis.ind_mat, ks, irs) * wl.conv_unit_e
    if whole:
        _ABS_CACHE[_wl_key(wl)] = A
    return A


def probe(dev, which, K, rng):
    from mqct_amd import _lib as L
    A = rng.standard_normal((16, K))
    B = rng.standard_normal((K, 16))
    if which == 2:
        A = A.astype(np.float32).astype(np.float64)
        B = B.astype(np.float32).astype(np.float64)
        # bf16-exact values so the product is exact up to fp32 summation
        A = np.round(A * 8) / 8
        B = np.round(B * 8) / 8
    C = np.zeros((16, 16))
    L.check(dev.lib.mqct_hip_probe_mfma(which, L.dptr(np.ascontiguousarray(A)), L.dptr(np.ascontiguousarray(B)),
                                        L.dptr(C)), "probe")
    return A @ B, C


@pytest.mark.parametrize("which,K,tol", [(0, 4, 1e-14), (1, 4, 1e-5), (2, 32, 1e-5)])
def test_mfma_lane_layout(dev, which, K, tol):
    ref, C = probe(dev, which, K, np.random.default_rng(which))
    assert np.max(np.abs(C - ref)) <= tol * np.max(np.abs(ref))


_ORACLE_CACHE = {}


def _wl_key(wl):
    return (wl.name, wl.basis.n_states, wl.basis.total_size, wl.grid.G, np.asarray(wl.R).tobytes(),
            np.asarray(wl.pes_par).tobytes(), wl.conv_unit_e, wl.mij_zero_cut, wl.ir_cut)


def build_both(dev, wl, precision="fp64"):
    """The GPU build and the oracle's (the oracle's result is kept per
    workload, so the variant tests of one workload run the oracle once)."""
    mb = MatrixBuilder(dev, wl)
    M, D, S = mb.build(precision)
    key = _wl_key(wl)
    if key not in _ORACLE_CACHE:
        orc = O.oracle_for_workload(wl)
        Mo, Do, So = orc.matrix_make(wl.basis.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
        _ORACLE_CACHE[key] = ((Mo, Do, So), orc)
    (Mo, Do, So), orc = _ORACLE_CACHE[key]
    return (M, D, S), (Mo.copy(), Do.copy(), So.copy()), orc


def test_diatom_atom_fp64_full_size(dev):
    """121 states, 506 pairs, 64 GL points, 200 R: every element."""
    wl = W.diatom_atom()
    (M, D, S), (Mo, Do, So), orc = build_both(dev, wl)
    assert np.array_equal(S, So)
    A = abs_sums(orc, wl)
    assert fp64_ok(M, Mo, A) <= 1.0
    assert rel_err(D, Do, 1e-10) <= 1e-8


def test_diatom_atom_fp64_long_grid(dev):
    """700 R points: the spline's scans run past the 512 points kept in
    registers, through the derivative table."""
    wl = W.diatom_atom(jmax=4, n_r=700)
    (M, D, S), (Mo, Do, So), orc = build_both(dev, wl)
    assert np.array_equal(S, So)
    A = abs_sums(orc, wl)
    assert fp64_ok(M, Mo, A) <= 1.0
    assert rel_err(D, Do, 1e-10) <= 1e-8


def test_asym_top_atom_fp64_small(dev):
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=20)
    (M, D, S), (Mo, Do, So), orc = build_both(dev, wl)
    assert np.array_equal(S, So)
    A = abs_sums(orc, wl)
    assert fp64_ok(M, Mo, A) <= 1.0
    assert rel_err(D, Do, 1e-10) <= 1e-8


def test_asym_top_atom_refxz_no(dev):
    """REF_XZ=NO integrand (pot_comp.f:2544-2553)."""
    wl = W.asym_top_atom(jmax=2, n_beta=8, n_gamma=16, n_r=12, ref_xz=0)
    (M, D, S), (Mo, Do, So), orc = build_both(dev, wl)
    assert np.array_equal(S, So)
    assert fp64_ok(M, Mo, abs_sums(orc, wl)) <= 1.0


@pytest.mark.parametrize("ref_xz", [1, 0])
def test_asym_top_atom_fp64_closed_form(dev, ref_xz):
    """The GPU's FP64 build (host V grid through mqct_hip_set_vgrid) of
    K-mixed asymmetric-top states (j <= 4) in a potential built from
    D^l_{0q} (l <= 4) against the 3j closed form of tests/closed_forms.py,
    in the XZ frame and with REF_XZ=NO (pot_comp.f:2544-2553): 1e-11 of the
    largest element (the 28 beta points converge to ~1e-14)."""
    import closed_forms as CF
    from mqct_amd import basis as B
    from mqct_amd.workloads import Workload
    A_, B_, C_ = W.ROT_A, W.ROT_B, W.ROT_C
    b = B.ini_arrays(B.channels_type4(0, 4, A_, B_, C_))
    g = B.grid_type4(28, 24, ref_xz=ref_xz)
    beta = g.xbm + g.xbl * np.asarray(g.xb)
    gamma = np.asarray(g.xg)
    rng = np.random.default_rng(43)
    R = np.array([4.0, 5.0, 6.0, 8.0])
    ws = [CF.d_potential_coeffs(rng, 4) for _ in R]
    V = np.stack([CF.d_potential_grid(w, beta, gamma) for w in ws])
    wl = Workload("d_potential", b, g, R, 1, np.zeros(1), 3.27, conv_unit_e=1.0, mij_cutoff_cm=0.0)
    M, D, S = MatrixBuilder(dev, wl, vgrid=V.ravel()).build("fp64")
    ref = np.zeros_like(M)
    for k in range(b.total_size):
        s1, s2 = b.ind_mat[k] - 1
        c1, c2 = b.indx_chann[s1] - 1, b.indx_chann[s2] - 1
        J1, J2, Mq = int(b.j1_ch[c1]), int(b.j1_ch[c2]), int(b.m12[s1])
        for ir in range(len(R)):
            ref[k, ir] = CF.type4_element(ws[ir], J1, b.eig1[c1, :2 * J1 + 1], J2, b.eig1[c2, :2 * J2 + 1], Mq,
                                          ref_xz=ref_xz)
    err = np.max(np.abs(M - ref)) / np.abs(ref).max()
    print(f"ref_xz {ref_xz}: {b.total_size} pairs x {len(R)} R, max err / max |M| {err:.2e}")
    assert err <= 1e-11


def test_asym_top_atom_matches_reference_mtrx_dat(dev):
    """The GPU's FP64 build of the reference's own SYS_TYPE 4 example
    (H2O + He, explicit K-mixed eigenvectors, the V grid of the example's
    own PES) against the example's MTRX.DAT: within 1e-12 of the largest
    element (12 printed digits), 1e-11 relative where significant, and the
    cutoff mask equal to the file's zero elements."""
    import closed_forms as CF
    from mqct_amd import basis as B
    from mqct_amd.workloads import Workload
    b, g, R, V, ref = CF.h2ohe_reference_system()
    wl = Workload("h2o_he", b, g, R, 1, np.zeros(1), 3.2748)
    M, D, S = MatrixBuilder(dev, wl, vgrid=V.ravel()).build("fp64")
    scale = np.abs(ref).max()
    err = np.max(np.abs(M - ref)) / scale
    print(f"H2O-He MTRX.DAT: max |gpu - file| / max |M| {err:.2e}")
    assert np.array_equal(S.astype(bool), np.all(ref == 0.0, axis=1))
    assert err <= 1e-12
    big = np.abs(ref) > 1e-3 * scale
    assert np.max(np.abs(M[big] - ref[big]) / np.abs(ref[big])) <= 1e-11


def test_two_asym_tops_fp64_small(dev):
    wl = W.two_asym_tops(jmax1=1, jmax2=1, n_r=8, n_ang_each=(2, 3, 3, 2, 3))
    (M, D, S), (Mo, Do, So), orc = build_both(dev, wl)
    assert np.array_equal(S, So)
    assert fp64_ok(M, Mo, abs_sums(orc, wl)) <= 1.0


@pytest.fixture(scope="module")
def c1_sample():
    """Config 1 at full size (1771 states, 86 779 pairs, 24x48 grid, 300 R):
    the oracle's values of a seeded sample of 150 elements at 5 R points,
    with their absolute-term sums A."""
    wl = W.asym_top_atom()
    orc = O.oracle_for_workload(wl)
    rng = np.random.default_rng(7)
    ks = rng.choice(wl.basis.total_size, 150, replace=False)
    irs = [0, 57, 149, 233, 299]
    A = abs_sums(orc, wl, ks, irs)
    Mo = np.array([[orc.element(wl.basis.ind_mat, int(k), ir) * wl.conv_unit_e for ir in irs] for k in ks])
    return wl, ks, irs, A, Mo


def test_asym_top_atom_full_size_sampled(dev, c1_sample):
    """FP64 at full size against the oracle on the sample; reports how many
    sampled values miss a pure 1e-10 relative bar (strong cancellation:
    |value| far below A) besides the combined bar."""
    wl, ks, irs, A, Mo = c1_sample
    M, D, S = MatrixBuilder(dev, wl).build("fp64")
    G = M[ks][:, irs]
    live = ~S[ks].astype(bool)
    worst = np.max((np.abs(G - Mo) / np.maximum(np.maximum(1e-10 * np.abs(Mo), 1e-13 * A), 1e-300))[live])
    rel = np.abs(G - Mo) / np.maximum(np.abs(Mo), 1e-300)
    miss = (rel > 1e-10) & live[:, None]
    print(f"FP64 full size: {int(miss.sum())} of {int(live.sum()) * len(irs)} sampled values miss 1e-10 relative; "
          f"their |value|/A is at most {np.max((np.abs(Mo) / A)[miss]) if miss.any() else 0:.1e}")
    assert worst <= 1.0
    if miss.any():
        assert np.max((np.abs(Mo) / A)[miss]) < 1e-3   # only strongly cancelling values


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_asym_top_atom_full_size_mixed_sampled(dev, c1_sample, prec):
    """Config 1's precision forms at full size against the oracle on the
    sample: the per-element maximum relative error (over the sampled R
    points, relative to the element's largest |value| there) is reported,
    and every value is held to 4e-6 A."""
    wl, ks, irs, A, Mo = c1_sample
    M, D, S = MatrixBuilder(dev, wl).build(prec)
    G = M[ks][:, irs]
    live = ~S[ks].astype(bool)
    per_el = (np.abs(G - Mo).max(axis=1) / np.maximum(np.abs(Mo).max(axis=1), 1e-300))[live]
    back = np.max((np.abs(G - Mo) / A)[live])
    print(f"{prec} full size: per-element max relative error {per_el.max():.2e} (median {np.median(per_el):.2e}), "
          f"|err|/A max {back:.2e}")
    assert back <= MIXED_TOL[prec]


# FP32: operands and chunk sums in FP32; split-BF16: three BF16 terms per
# operand (about 24 bits), six cross products; both with FP64 chunk totals.
MIXED_TOL = {"fp32": 4e-6, "bf16x3": 4e-6}


def check_mixed(dev, wl, prec):
    (M, D, S), (Mo, Do, So), orc = build_both(dev, wl, prec)
    assert np.array_equal(S, So)  # the mask comes from FP64 sums
    live = ~So.astype(bool)
    A = abs_sums(orc, wl)
    back = np.max((np.abs(M - Mo) / np.maximum(A, 1e-300))[live])
    colmax = np.abs(Mo).max(axis=1, keepdims=True)
    rel_col = np.max((np.abs(M - Mo) / np.maximum(colmax, 1e-300))[live])
    rel_glob = np.max(np.abs(M - Mo)) / np.abs(Mo).max()
    print(f"{wl.name} {prec}: |err|/A max {back:.3e}, |err|/colmax max {rel_col:.3e}, "
          f"|err|/max|M| {rel_glob:.3e}")
    assert back <= MIXED_TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_diatom_atom(dev, prec):
    check_mixed(dev, W.diatom_atom(), prec)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_odd_grid(dev, prec):
    """An odd number of grid points (63): the contractions take their forms
    without 16-byte staging."""
    check_mixed(dev, W.diatom_atom(jmax=5, n_beta=63, n_r=30), prec)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_asym_top(dev, prec):
    check_mixed(dev, W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=20), prec)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_two_asym_tops(dev, prec):
    """Type 0 with a grid of 192 points (a multiple of the 16-byte staging
    width of every form)."""
    check_mixed(dev, W.two_asym_tops(jmax1=1, jmax2=1, n_ang_each=(2, 4, 3, 2, 4), n_r=8), prec)



@pytest.mark.parametrize("prec,variant", [("fp64", 0), ("fp64", 1), ("fp64", 2), ("fp32", 0), ("fp32", 1),
                                          ("fp32", 2), ("fp32", 3), ("bf16x3", 0), ("bf16x3", 1), ("bf16x3", 2),
                                          ("bf16x3", 3)])
def test_contraction_variants(dev, prec, variant):
    """Every contraction variant the build may choose (grid points per LDS
    stage, waves per workgroup) meets its precision's bar against the
    oracle, on an asymmetric top whose grid (288 points) suits the 16-byte
    staging of every form, and 150 R points (R blocks of 64 and 128 with a
    partial last block); the timed choice is restored afterwards."""
    from mqct_amd import _lib as L
    from mqct_amd.matrix import PRECISIONS
    import ctypes
    p = PRECISIONS[prec]
    L.check(dev.lib.mqct_hip_set_contract_variant(dev.h, p, variant), "set_contract_variant")
    try:
        wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=150)
        if prec == "fp64":
            (M, D, S), (Mo, Do, So), orc = build_both(dev, wl)
            assert np.array_equal(S, So)
            assert fp64_ok(M, Mo, abs_sums(orc, wl)) <= 1.0
        else:
            check_mixed(dev, wl, prec)
        v, kc, nw, nc = (ctypes.c_int() for _ in range(4))
        L.check(dev.lib.mqct_hip_contract_variant(dev.h, p, ctypes.byref(v), ctypes.byref(kc), ctypes.byref(nw),
                                                  ctypes.byref(nc)), "contract_variant")
        assert v.value == variant and kc.value > 0 and nw.value in (4, 8) and nc.value > variant
    finally:
        L.check(dev.lib.mqct_hip_set_contract_variant(dev.h, p, -1), "set_contract_variant")


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_unaligned_grid(dev, prec):
    """A grid of 45 points (5 x 9), not a multiple of the 16-byte staging
    width: the builds take the scalar-staging kernels."""
    check_mixed(dev, W.asym_top_atom(jmax=2, n_beta=5, n_gamma=9, n_r=20), prec)
