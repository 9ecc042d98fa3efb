"""ctypes front end of the CPU oracle (oracle/_build/libmqct_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always to check against or as the timed CPU
baseline.  The product (mqct_amd) never imports it.

Each wrapper names the oracle C function it calls; those name the reference
lines they restate.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmqct_oracle.so")

_i = ctypes.c_int
_d = ctypes.c_double
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class MatrixCtx(ctypes.Structure):
    _fields_ = [
        ("coll_type", _i), ("n_states", _i),
        ("j12", _ip), ("m12", _ip), ("indx_chann", _ip),
        ("j1_ch", _ip), ("j2_ch", _ip),
        ("ldvec", _i), ("eig1", _dp), ("eig2", _dp),
        ("ref_xz", _i),
        ("n_beta", _i), ("xb", _dp), ("wb", _dp), ("xbm", _d), ("xbl", _d),
        ("n_gamma", _i), ("xg", _dp), ("wg", _dp), ("xgm", _d), ("xgl", _d),
        ("n_alpha2", _i), ("xaa", _dp), ("waa", _dp), ("xaam", _d), ("xaal", _d),
        ("n_beta2", _i), ("xbb", _dp), ("wbb", _dp), ("xbbm", _d), ("xbbl", _d),
        ("n_gamma2", _i), ("xgg", _dp), ("wgg", _dp), ("xggm", _d), ("xggl", _d),
        ("eq_gamma", _i),
        ("n_r", _i),
        ("V", _dp),
        ("abs_mode", _i),
        ("identical", _i), ("identical_pes", _i),
        ("parity_state", _ip), ("parity_state_bk", _ip),
        ("ka1_ch", _ip), ("kc1_ch", _ip), ("ka2_ch", _ip), ("kc2_ch", _ip),
        ("plant", _i),
    ]


class DynCtx(ctypes.Structure):
    _fields_ = [
        ("n_states", _i), ("total_size", ctypes.c_int64), ("ind_mat", _ip),
        ("j12", _ip), ("m12", _ip), ("Ej", _dp),
        ("n_r", _i), ("R", _dp), ("mat", _dp), ("mat_der", _dp),
        ("mass", _d), ("theta_on", _i),
        ("misc", _dp), ("bk_indx", _ip), ("delta_E", _dp), ("n_dE", _i),
        ("at_n", _i), ("at_t", _dp), ("at_var", _dp), ("at_var_der", _dp),
        ("plant", _i),
    ]


class TrajParam(ctypes.Structure):
    _fields_ = [
        ("mode", _i), ("s_st", _i), ("l_real", _d), ("E_coll", _d), ("dt", _d),
        ("eps", _d), ("h1", _d), ("hmin", _d), ("R_start", _d), ("time_lim", _d),
        ("numb_oscl_prds", _i), ("numb_orb_prds", _i), ("max_steps", _i),
    ]


class TrajStat(ctypes.Structure):
    _fields_ = [("t_end", _d), ("n_steps", _i), ("n_rhs", _i), ("fail", _i), ("norm_err", _d)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.or_jacobi_poly.argtypes = [_i, _d, _d, _d, _dp]
        L.or_djkm.argtypes = [_i, _i, _i, _d, _d, _d, _dp, _dp]
        L.or_plgndr.argtypes = [_i, _i, _d]
        L.or_plgndr.restype = _d
        L.or_w3js.argtypes = [_i] * 6
        L.or_w3js.restype = _d
        L.or_cg.argtypes = [_i] * 6
        L.or_cg.restype = _d
        L.or_gauleg.argtypes = [_i, _dp, _dp]
        L.or_wf_diat_top.argtypes = [_i, _i, _d, _d, _dp, _dp]
        L.or_wf_asym_top.argtypes = [_i, _dp, _i, _d, _d, _d, _dp, _dp]
        L.or_hamilt_kyro.argtypes = [_i, _i, _i, _i, _d, _d, _d]
        L.or_hamilt_kyro.restype = _d
        L.or_state.argtypes = [_i, _i, _ip, _ip]
        L.or_asym_top_eigen.argtypes = [_d, _d, _d, _i, _dp, _dp]
        L.or_eigen_vec.argtypes = [_d, _d, _d, _i, _i, _i, _dp, _dp]
        L.or_channels_type1.argtypes = [_i, _i, _d, _d, _ip, _dp]
        L.or_channels_type4.argtypes = [_i, _i, _d, _d, _d, _ip, _ip, _ip, _dp]
        L.or_channels_type0.argtypes = [_i, _i, _i, _i, _d, _d, _d, _d, _d, _d] + [_ip] * 6 + [_dp]
        L.or_ini_arrays.argtypes = [_i, _i, _ip, _ip, _ip, _ip, _ip, _ip, ctypes.c_int64, _ip, _ip]
        L.or_ini_arrays.restype = ctypes.c_int64
        L.or_integrator_temp.argtypes = [ctypes.POINTER(MatrixCtx), _ip, ctypes.c_int64, _i]
        L.or_integrator_temp.restype = _d
        L.or_matrix_make.argtypes = [ctypes.POINTER(MatrixCtx), ctypes.c_int64, _ip, _dp, _d, _d, _i, _i,
                                     _dp, _dp, _u8p]
        L.or_spline.argtypes = [_dp, _dp, _i, _d, _d, _dp]
        L.or_mij_shift.argtypes = [ctypes.c_int64, _i, _dp]
        L.or_read_mij_post.argtypes = [_dp, _i, ctypes.c_int64, _i, _dp, _dp]
        L.or_splint.argtypes = [_dp, _dp, _dp, _i, _d, _dp, _dp]
        L.or_nearest_r.argtypes = [_dp, _i, _d]
        L.or_pes_legendre.argtypes = [_dp, _d, _d]
        L.or_pes_legendre.restype = _d
        L.or_pes_realsh.argtypes = [_dp, _d, _d, _d]
        L.or_pes_realsh.restype = _d
        L.or_pes_sites.argtypes = [_dp, _d, _d, _d, _d, _d, _d]
        L.or_pes_sites.restype = _d
        L.or_vgrid.argtypes = [_i, _i, _dp, ctypes.POINTER(MatrixCtx), _dp, _dp]
        L.or_expansion_term.argtypes = [_i, _ip, _d, _d, _d, _d, _d, _dp, _dp]
        L.or_expansion_term_grid.argtypes = [ctypes.POINTER(MatrixCtx), _ip, _dp]
        L.or_calc_expansion_terms.argtypes = [ctypes.POINTER(MatrixCtx), _i, _ip, _dp]
        L.or_expansion_matrix_element.argtypes = [ctypes.POINTER(MatrixCtx), _i, _ip, _dp, _ip, ctypes.c_int64,
                                                  _dp]
        L.or_expansion_matrix_make.argtypes = [ctypes.POINTER(MatrixCtx), _i, _ip, _dp, ctypes.c_int64, _ip, _dp,
                                               _d, _d, _i, _i, _dp, _dp, _u8p]
        L.or_logfact_init.argtypes = [_i]
        L.or_resize.argtypes = [ctypes.POINTER(DynCtx)]
        L.or_at_prepare.argtypes = [ctypes.POINTER(DynCtx)]
        L.or_potential_data_bk.argtypes = [ctypes.POINTER(DynCtx), _d, _dp, _dp]
        L.or_derivs_bk.argtypes = [ctypes.POINTER(DynCtx), _d, _dp, _dp]
        L.or_derivs_bk_adia.argtypes = [ctypes.POINTER(DynCtx), _d, _dp, _dp]
        L.or_dotraject.argtypes = [ctypes.POINTER(DynCtx), ctypes.POINTER(TrajParam), _dp,
                                   ctypes.POINTER(TrajStat), _dp, _i, _ip]
        L.or_logfact_init(400)
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def dp(a):
    return _p(a, ctypes.c_double)


def ip(a):
    return _p(a, ctypes.c_int)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------------------
def gauleg(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_gauleg(n, dp(x), dp(w))
    return x, w


def djkm(j, k, m, alpha, beta, gamma):
    re, im = _d(), _d()
    lib().or_djkm(j, k, m, alpha, beta, gamma, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def wf_diat_top(j, m, theta, phi):
    re, im = _d(), _d()
    lib().or_wf_diat_top(j, m, theta, phi, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def wf_asym_top(J, coeff, M, alpha, beta, gamma):
    c = f64(coeff)
    re, im = _d(), _d()
    lib().or_wf_asym_top(J, dp(c), M, alpha, beta, gamma, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def w3js(*a):
    return lib().or_w3js(*a)


def cg(*a):
    return lib().or_cg(*a)


def plgndr(l, m, x):
    return lib().or_plgndr(l, m, x)


def state(J, DK):
    ka, kc = _i(), _i()
    ok = lib().or_state(J, DK, ctypes.byref(ka), ctypes.byref(kc))
    return (ka.value, kc.value) if ok else None


def asym_top_eigen(A, B, C, J):
    n = 2 * J + 1
    w, z = np.zeros(n), np.zeros(n * n)
    lib().or_asym_top_eigen(A, B, C, J, dp(w), dp(z))
    return w, z.reshape(n, n).T  # columns = eigenvectors


def eigen_vec(A, B, C, J, KA, KC):
    E = _d()
    z = np.zeros(2 * J + 1)
    ok = lib().or_eigen_vec(A, B, C, J, KA, KC, ctypes.byref(E), dp(z))
    if not ok:
        raise ValueError("no such level")
    return E.value, z


def channels_type1(jmin, jmax, Be, De=0.0):
    n = jmax - max(jmin, 0) + 1
    j, E = np.zeros(n, np.int32), np.zeros(n)
    lib().or_channels_type1(jmin, jmax, Be, De, ip(j), dp(E))
    return {"j": j, "E": E}


def channels_type4(jmin, jmax, A, B, C):
    jmin = max(jmin, 0)
    n = (jmax - jmin + 1) * (jmax + jmin + 1)
    j, ka, kc = (np.zeros(n, np.int32) for _ in range(3))
    E = np.zeros(n)
    lib().or_channels_type4(jmin, jmax, A, B, C, ip(j), ip(ka), ip(kc), dp(E))
    return {"j": j, "ka": ka, "kc": kc, "E": E}


def channels_type0(jmin1, jmax1, jmin2, jmax2, A1, B1, C1, A2, B2, C2):
    n = (jmax1 - jmin1 + 1) * (jmax1 + jmin1 + 1) * (jmax2 - jmin2 + 1) * (jmax2 + jmin2 + 1)
    arr = [np.zeros(n, np.int32) for _ in range(6)]
    E = np.zeros(n)
    lib().or_channels_type0(jmin1, jmax1, jmin2, jmax2, A1, B1, C1, A2, B2, C2, *[ip(a) for a in arr], dp(E))
    return {"j1": arr[0], "ka1": arr[1], "kc1": arr[2], "j2": arr[3], "ka2": arr[4], "kc2": arr[5], "E": E}


def ini_arrays(coll_type, j1_ch, j2_ch=None):
    j1 = i32(j1_ch)
    j2 = i32(j2_ch) if j2_ch is not None else None
    ns = _i()
    lib().or_ini_arrays(coll_type, len(j1), ip(j1), ip(j2), None, None, None, None, 0, None, ctypes.byref(ns))
    S = ns.value
    j12, m12, ich = (np.zeros(S, np.int32) for _ in range(3))
    cidx = np.zeros(len(j1), np.int32)
    T = lib().or_ini_arrays(coll_type, len(j1), ip(j1), ip(j2), ip(j12), ip(m12), ip(ich), ip(cidx), 0, None,
                            None)
    ind = np.zeros((T, 2), np.int32)
    lib().or_ini_arrays(coll_type, len(j1), ip(j1), ip(j2), ip(j12), ip(m12), ip(ich), ip(cidx), T, ip(ind),
                        None)
    return {"j12": j12, "m12": m12, "indx_chann": ich, "chann_indx": cidx, "ind_mat": ind}


class MatrixOracle:
    """INTEGRATOR_TEMP / MATRIX_MAKE on the CPU for a system given as plain
    arrays (basis tables, grid, V grid)."""

    def __init__(self, coll_type, j12, m12, indx_chann, j1_ch, grid, R, V, eig1=None, eig2=None, j2_ch=None):
        self.keep = []

        def k(a, f):
            if a is None:
                return None
            a = f(a)
            self.keep.append(a)
            return a

        c = MatrixCtx()
        c.coll_type = coll_type
        c.n_states = len(j12)
        c.j12, c.m12, c.indx_chann = ip(k(j12, i32)), ip(k(m12, i32)), ip(k(indx_chann, i32))
        c.j1_ch = ip(k(j1_ch, i32))
        c.j2_ch = ip(k(j2_ch, i32))
        if eig1 is not None:
            c.ldvec = eig1.shape[1]
        c.eig1, c.eig2 = dp(k(eig1, f64)), dp(k(eig2, f64))
        c.ref_xz = grid.get("ref_xz", 1)
        one = np.ones(1)
        zero = np.zeros(1)
        for nm, pre in (("beta", "b"), ("gamma", "g"), ("alpha2", "aa"), ("beta2", "bb"), ("gamma2", "gg")):
            n = grid.get("n_" + nm, 1)
            setattr(c, "n_" + nm, n)
            setattr(c, "x" + pre, dp(k(grid.get("x" + pre, zero), f64)))
            setattr(c, "w" + pre, dp(k(grid.get("w" + pre, one), f64)))
            setattr(c, "x" + pre + "m", grid.get("x" + pre + "m", 0.0))
            setattr(c, "x" + pre + "l", grid.get("x" + pre + "l", 1.0))
        c.eq_gamma = grid.get("eq_gamma", 1)
        c.n_r = len(R)
        self.R = k(R, f64)
        self.V = k(V, f64)
        c.V = dp(self.V)
        self.c = c

    def set_identical(self, basis, identical=True, identical_pes=True):
        """IDENTICAL= / IDENTICAL_PES= of a type-0 basis (mqct_amd.basis)."""
        c = self.c
        c.identical, c.identical_pes = int(identical), int(identical_pes or identical)
        keep = lambda a: ip(self.keep.append(i32(a)) or self.keep[-1])
        if identical:
            c.parity_state = keep(basis.parity_state)
            c.parity_state_bk = keep(basis.parity_state_bk)
        c.ka1_ch, c.kc1_ch = keep(basis.ka1_ch), keep(basis.kc1_ch)
        c.ka2_ch, c.kc2_ch = keep(basis.ka2_ch), keep(basis.kc2_ch)

    def element(self, ind_mat, k, i_r):
        ind = i32(ind_mat)
        return lib().or_integrator_temp(ctypes.byref(self.c), ip(ind), k, i_r)

    def abs_sums(self, ind_mat, ks, irs):
        """sum_g |integrand| of elements ks at R points irs: the scale the
        rounding error of any summation order is proportional to."""
        ind = i32(ind_mat)
        self.c.abs_mode = 1
        try:
            return np.array([[lib().or_integrator_temp(ctypes.byref(self.c), ip(ind), int(k), int(i))
                              for i in irs] for k in ks])
        finally:
            self.c.abs_mode = 0

    def matrix_make(self, ind_mat, conv_unit_e, mij_zero_cut, ir1, ir2):
        ind = i32(ind_mat)
        T, n = len(ind), len(self.R)
        M, D = np.zeros((T, n)), np.zeros((T, n))
        S = np.zeros(T, np.uint8)
        lib().or_matrix_make(ctypes.byref(self.c), T, ip(ind), dp(self.R), conv_unit_e, mij_zero_cut, ir1, ir2,
                             dp(M), dp(D), S.ctypes.data_as(_u8p))
        return M, D, S


    # -- PES expansion (mqct_oracle_exp.c) --------------------------------
    def expansion_terms(self, a_top):
        """CALC_EXPANSION_TERMS: [n_terms][n_r] coefficients of the V grid."""
        a = _a_top(a_top)
        out = np.zeros((len(a), len(self.R)))
        lib().or_calc_expansion_terms(ctypes.byref(self.c), len(a), ip(a), dp(out))
        return out

    def term_grid(self, term):
        """Real part of the term's function on the grid (V-grid order)."""
        c = self.c
        G = c.n_beta * c.n_gamma * c.n_alpha2 * c.n_beta2 * c.n_gamma2
        out = np.zeros(G)
        t = _a_top([term])
        lib().or_expansion_term_grid(ctypes.byref(c), ip(t), dp(out))
        return out

    def expansion_element(self, a_top, terms, ind_mat, k):
        """EXPANSION_MATRIX_ELEMENT of pair k (0-based) at every R point."""
        a, t, ind = _a_top(a_top), f64(terms), i32(ind_mat)
        out = np.zeros(len(self.R))
        lib().or_expansion_matrix_element(ctypes.byref(self.c), len(a), ip(a), dp(t), ip(ind), int(k), dp(out))
        return out

    def expansion_matrix_make(self, a_top, terms, ind_mat, conv_unit_e, mij_zero_cut, ir1, ir2):
        """MATRIX_MAKE with expansion_defined: M, Mat_el_der, skipped."""
        a, t, ind = _a_top(a_top), f64(terms), i32(ind_mat)
        T, n = len(ind), len(self.R)
        M, D = np.zeros((T, n)), np.zeros((T, n))
        S = np.zeros(T, np.uint8)
        lib().or_expansion_matrix_make(ctypes.byref(self.c), len(a), ip(a), dp(t), T, ip(ind), dp(self.R),
                                       conv_unit_e, mij_zero_cut, ir1, ir2, dp(M), dp(D), S.ctypes.data_as(_u8p))
        return M, D, S


def _a_top(a_top):
    """Term rows padded to A_TOP's five slots."""
    a = np.zeros((len(a_top), 5), np.int32)
    for i, row in enumerate(a_top):
        a[i, :len(row)] = row
    return a


def expansion_term(coll_type, term, b1, g1=0.0, a2=0.0, b2=0.0, g2=0.0):
    t = _a_top([term])
    re, im = _d(), _d()
    lib().or_expansion_term(coll_type, ip(t), b1, g1, a2, b2, g2, ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


def vgrid(coll_type, pes_kind, par, ctx: MatrixOracle, R):
    c = ctx.c
    G = c.n_beta * c.n_gamma * c.n_alpha2 * c.n_beta2 * c.n_gamma2
    V = np.zeros(len(R) * G)
    p = f64(par)
    lib().or_vgrid(coll_type, pes_kind, dp(p), ctypes.byref(c), dp(f64(R)), dp(V))
    return V


def spline(x, y, yp1, ypn):
    x, y = f64(x), f64(y)
    y2 = np.zeros(len(x))
    lib().or_spline(dp(x), dp(y), len(x), yp1, ypn, dp(y2))
    return y2


def mij_shift(mat_el):
    """or_mij_shift: the compute branch's MIJ_SHIFT (matrix_ini.f:2226-2230)
    on a copy of mat_el[k, i]."""
    M = np.array(mat_el, dtype=np.float64, order="C", copy=True)
    lib().or_mij_shift(M.shape[0], M.shape[1], dp(M))
    return M


def read_mij_post(R, mat_el, shift=True):
    """or_read_mij_post: READ_MIJ_USER's shift and re-spline
    (matrix_ini.f:6603-6621) on a copy of mat_el[k, i]; returns (M, D)."""
    M = np.array(mat_el, dtype=np.float64, order="C", copy=True)
    D = np.zeros_like(M)
    R = f64(R)
    lib().or_read_mij_post(dp(R), len(R), M.shape[0], int(shift), dp(M), dp(D))
    return M, D


def splint(xa, ya, y2a, x):
    xa, ya, y2a = f64(xa), f64(ya), f64(y2a)
    y, yp = _d(), _d()
    lib().or_splint(dp(xa), dp(ya), dp(y2a), len(xa), x, ctypes.byref(y), ctypes.byref(yp))
    return y.value, yp.value


def grid_dict(g):
    """AngularGrid (mqct_amd.basis) or dict -> oracle grid dict."""
    if isinstance(g, dict):
        return g
    d = {}
    for nm, pre in (("beta", "b"), ("gamma", "g"), ("alpha2", "aa"), ("beta2", "bb"), ("gamma2", "gg")):
        d["n_" + nm] = getattr(g, "n_" + nm)
        for s in ("x" + pre, "w" + pre):
            v = getattr(g, s)
            if v is not None:
                d[s] = v
        d["x" + pre + "m"] = getattr(g, "x" + pre + "m")
        d["x" + pre + "l"] = getattr(g, "x" + pre + "l")
    d["eq_gamma"] = g.eq_gamma
    d["ref_xz"] = g.ref_xz
    return d


def oracle_for_workload(wl, vgrid_values=None):
    """MatrixOracle for a workloads.Workload, with the V grid from the
    oracle's own potential unless given."""
    b = wl.basis
    g = grid_dict(wl.grid)
    o = MatrixOracle(b.coll_type, b.j12, b.m12, b.indx_chann, b.j1_ch, g, wl.R, np.zeros(1),
                     eig1=b.eig1, eig2=b.eig2, j2_ch=b.j2_ch)
    if vgrid_values is None:
        vgrid_values = vgrid(b.coll_type, wl.pes_kind, wl.pes_par, o, wl.R)
    o.V = f64(vgrid_values)
    o.keep.append(o.V)
    o.c.V = dp(o.V)
    return o


class DynOracle:
    """DERIVS_BK / DERIVS_BK_adia and the DOTRAJECT loop on the CPU
    (or_dyn_ctx of mqct_oracle_dyn.c)."""

    def __init__(self, j12, m12, Ej_hartree, ind_mat, R, mat, mat_der, mass, theta_on=0, path=None):
        self.keep = []

        def k(a, f):
            a = f(a)
            self.keep.append(a)
            return a

        d = DynCtx()
        self.N = len(j12)
        T = len(ind_mat)
        d.n_states = self.N
        d.total_size = T
        d.ind_mat = ip(k(ind_mat, i32))
        d.j12, d.m12 = ip(k(j12, i32)), ip(k(m12, i32))
        d.Ej = dp(k(Ej_hartree, f64))
        d.n_r = len(R)
        self.R = k(R, f64)
        d.R = dp(self.R)
        d.mat, d.mat_der = dp(k(mat, f64)), dp(k(mat_der, f64))
        d.mass = mass
        d.theta_on = theta_on
        self.misc = np.zeros(4 * T)
        self.bk_indx = np.zeros(T, np.int32)
        self.delta_E = np.zeros(T)
        d.misc, d.bk_indx, d.delta_E = dp(self.misc), ip(self.bk_indx), dp(self.delta_E)
        self.d = d
        lib().or_resize(ctypes.byref(d))
        if path is not None:
            self.set_path(path)

    def set_path(self, rec):
        """rec: [n][8] records (t, R, P_R, Th, P_Th, Phi, P_Phi, V)."""
        rec = np.asarray(rec)
        self.at_t = f64(rec[:, 0])
        self.at_var = f64(rec[:, 1:8].T)  # [7][n]
        self.at_der = np.zeros_like(self.at_var)
        d = self.d
        d.at_n = len(rec)
        d.at_t, d.at_var, d.at_var_der = dp(self.at_t), dp(self.at_var), dp(self.at_der)
        lib().or_at_prepare(ctypes.byref(d))

    def potential(self, R):
        """or_potential_data_bk: M_k(R) and dM_k/dR of every pair."""
        m, dm = np.zeros(self.d.total_size), np.zeros(self.d.total_size)
        lib().or_potential_data_bk(ctypes.byref(self.d), R, dp(m), dp(dm))
        return m, dm

    def derivs(self, t, q, adia=False):
        q = f64(q)
        dq = np.zeros(2 * self.N + 8)
        fn = lib().or_derivs_bk_adia if adia else lib().or_derivs_bk
        fn(ctypes.byref(self.d), t, dp(q), dp(dq))
        return dq

    def dotraject(self, mode, s_st, l_real, E_coll, dt, eps=1e-8, hmin=1e-6, R_start=None,
                  numb_oscl_prds=1, numb_orb_prds=1, max_steps=0, rec_cap=0):
        R_start = float(self.R[-1]) if R_start is None else R_start  # RMAX = last R point
        p = TrajParam(mode=mode, s_st=s_st, l_real=l_real, E_coll=E_coll, dt=dt, eps=eps, h1=dt, hmin=hmin,
                      R_start=R_start, time_lim=-1.0, numb_oscl_prds=numb_oscl_prds,
                      numb_orb_prds=numb_orb_prds, max_steps=max_steps)
        q = np.zeros(2 * self.N + 8)
        st = TrajStat()
        rec = np.zeros((max(rec_cap, 1), 8))
        rn = _i()
        lib().or_dotraject(ctypes.byref(self.d), ctypes.byref(p), dp(q), ctypes.byref(st), dp(rec), rec_cap,
                           ctypes.byref(rn))
        out = {"t_end": st.t_end, "n_steps": st.n_steps, "n_rhs": st.n_rhs, "fail": st.fail,
               "norm_err": st.norm_err}
        return q, out, rec[:rn.value].copy()
