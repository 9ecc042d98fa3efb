"""Host-side basis set-up: channel lists, asymmetric-top eigenvectors, state
tables and the pair list, as MQCT's driver builds them before the matrix
build.

Mirrors (names follow the reference):
  channels_type1 / _type4 / _type0 : OUTPUT channel generation from
      JMIN/JMAX (iotest.f:5426-5440, 6600-6626, 10321-10411)
  asym_top_eigen / eigen_vec        : EIGEN_VAL / EIGEN_VEC (asym_top.f)
  ini_arrays                        : INI_ARRAYS (matrix_ini.f:1-640)
  gauss_legendre / angle grids      : gauleg + INI_POT_STORAGE grid maps
                                      (pot_comp.f:374-382, 586-611, 1628-1709)
  r_grid                            : MATRIX_MAKE R grid (matrix_ini.f:954-1033)

Everything here is set-up work the Fortran driver already does on the host;
the HIP library receives its results.
"""
from dataclasses import dataclass, field

import numpy as np

PI = float(np.arccos(-1.0))
AUTOEV = 27.21113845          # iotest.f:20
EVTOWN = 8065.5448            # iotest.f:23
AMUTOAU = 1822.8885           # iotest.f:23
A_BOHR = 0.52917721067        # iotest.f:19
CM_TO_HARTREE = 1.0 / EVTOWN / AUTOEV


# ---------------------------------------------------------------------------
# asymmetric top
def state_labels(J, DK):
    """(Ka, Kc) of the (J+DK+1)-th eigenvalue of level J (asym_top.f:139-173)."""
    ka = [0] * (2 * J + 1)
    for i in range(2, 2 * J + 1, 2):
        ka[i - 1] = i // 2
        ka[i] = i // 2
    out = None
    for i in range(1, 2 * J + 2):
        kc = ka[2 * J + 2 - i - 1]
        if ka[i - 1] - kc == DK:
            out = (ka[i - 1], kc)
    return out


def rotor_matrix(A, B, C, J):
    """Rigid-rotor matrix in |J K>, K=-J..J (asym_top.f:103-137)."""
    n = 2 * J + 1
    H = np.zeros((n, n))
    jj = J * (J + 1.0)
    for K in range(-J, J + 1):
        H[K + J, K + J] = 0.5 * (A + B) * jj + (C - 0.5 * (A + B)) * K * K
        for K2 in (K - 2, K + 2):
            if -J <= K2 <= J:
                kav = (K + K2) // 2
                H[K + J, K2 + J] = 0.5 * (A - B) * (
                    np.sqrt((J + 1.0) * J - kav * K) * np.sqrt((J + 1.0) * J - kav * K2) / 2.0)
    return H


def asym_top_eigen(A, B, C, J):
    """Eigenvalues ascending and eigenvectors (columns) of the rotor matrix.

    The matrix splits exactly into four Wang blocks (K even/odd x symmetric/
    antisymmetric under K -> -K); each is diagonalised with LAPACK (numpy
    eigh), so nearly degenerate levels of different symmetry never mix.
    Sign convention: the largest |c_K| (lowest K on ties) is positive.
    """
    n = 2 * J + 1
    H = rotor_matrix(A, B, C, J)
    vals, vecs = [], []
    for par in (0, 1):
        for sym in (0, 1):
            basis = []
            for K in range(0, J + 1):
                if K % 2 != par:
                    continue
                u = np.zeros(n)
                if K == 0:
                    if sym == 0:
                        u[J] = 1.0
                        basis.append(u)
                    continue
                u[J + K] = 1.0 / np.sqrt(2.0)
                u[J - K] = (1.0 if sym == 0 else -1.0) / np.sqrt(2.0)
                basis.append(u)
            if not basis:
                continue
            U = np.array(basis)
            w, v = np.linalg.eigh(U @ H @ U.T)
            for c in range(len(w)):
                vals.append(w[c])
                vecs.append(U.T @ v[:, c])
    order = np.argsort(np.array(vals), kind="stable")
    w = np.array(vals)[order]
    z = np.array(vecs)[order].T
    for c in range(n):
        big = np.max(np.abs(z[:, c]))
        piv = int(np.nonzero(np.abs(z[:, c]) >= big * (1 - 1e-9))[0][0])
        if z[piv, c] < 0:
            z[:, c] = -z[:, c]
    return w, z


def eigen_vec(A, B, C, J, KA, KC):
    """(E, c_K) of level (J, KA, KC): EIGEN_VAL/EIGEN_VEC (asym_top.f)."""
    w, z = asym_top_eigen(A, B, C, J)
    for K1 in range(-J, J + 1):
        if state_labels(J, K1) == (KA, KC):
            return w[J + K1], z[:, J + K1].copy()
    raise ValueError(f"no level J={J} Ka={KA} Kc={KC}")


# ---------------------------------------------------------------------------
@dataclass
class Basis:
    coll_type: int
    j1_ch: np.ndarray                 # per channel (j for types 1, 4)
    E_ch: np.ndarray                  # cm^-1
    ka1_ch: np.ndarray = None
    kc1_ch: np.ndarray = None
    j2_ch: np.ndarray = None
    ka2_ch: np.ndarray = None
    kc2_ch: np.ndarray = None
    eig1: np.ndarray = None           # [n_ch][ldvec]
    eig2: np.ndarray = None
    # state tables
    j12: np.ndarray = None
    m12: np.ndarray = None
    indx_chann: np.ndarray = None     # 1-based
    chann_indx: np.ndarray = None
    ind_mat: np.ndarray = None        # [T][2], 1-based
    Ej: np.ndarray = None             # cm^-1 per state
    # identical partners (type 0, IDENTICAL=YES): exchange parity per state
    # and the parity_state_bk sign the expansion's exchange terms carry
    identical: bool = False
    parity_state: np.ndarray = None
    parity_state_bk: np.ndarray = None
    extra: dict = field(default_factory=dict)

    @property
    def n_channels(self):
        return len(self.j1_ch)

    @property
    def n_states(self):
        return len(self.j12)

    @property
    def total_size(self):
        return len(self.ind_mat)

    @property
    def ldvec(self):
        return 0 if self.eig1 is None else self.eig1.shape[1]

    def chan_qn(self):
        """Per-channel quantum numbers of the MTRX_UF.DAT state records."""
        if self.coll_type == 1:
            return np.ascontiguousarray(self.j1_ch[:, None], dtype=np.int32)
        if self.coll_type == 4:
            return np.ascontiguousarray(np.stack([self.j1_ch, self.ka1_ch, self.kc1_ch], 1), dtype=np.int32)
        return np.ascontiguousarray(np.stack([self.j1_ch, self.ka1_ch, self.kc1_ch,
                                              self.j2_ch, self.ka2_ch, self.kc2_ch], 1), dtype=np.int32)

    def m_blocks(self):
        out = {}
        for s, m in enumerate(self.m12):
            out.setdefault(int(m), []).append(s)
        return out


def channels_type1(jmin, jmax, Be, De=0.0):
    jmin = max(jmin, 0)
    j = np.arange(jmin, jmax + 1)
    E = Be * j * (j + 1.0) - De * (j * (j + 1.0)) ** 2
    return Basis(1, j.astype(np.int32), E)


def channels_type4(jmin, jmax, A, B, C):
    jmin = max(jmin, 0)
    js, kas, kcs, Es, vecs = [], [], [], [], []
    ld = 2 * jmax + 1
    for j in range(jmin, jmax + 1):
        w, z = asym_top_eigen(A, B, C, j)
        for dk in range(-j, j + 1):
            ka, kc = state_labels(j, dk)
            js.append(j)
            kas.append(ka)
            kcs.append(kc)
            Es.append(w[j + dk])
            v = np.zeros(ld)
            v[:2 * j + 1] = z[:, j + dk]
            vecs.append(v)
    return Basis(4, np.array(js, np.int32), np.array(Es), np.array(kas, np.int32),
                 np.array(kcs, np.int32), eig1=np.array(vecs))


def _levels(jmin, jmax, A, B, C, ld):
    out = []
    for j in range(jmin, jmax + 1):
        w, z = asym_top_eigen(A, B, C, j)
        for dk in range(-j, j + 1):
            ka, kc = state_labels(j, dk)
            v = np.zeros(ld)
            v[:2 * j + 1] = z[:, j + dk]
            out.append((j, ka, kc, w[j + dk], v))
    return out


def channels_type0(jmin1, jmax1, jmin2, jmax2, A1, B1, C1, A2, B2, C2, keep=None, emax=None):
    """Distinguishable asym. top + asym. top; channel = (level1, level2),
    molecule 1 outermost (iotest.f:10383-10411).  keep(j1,ka1,kc1,j2,ka2,kc2)
    optionally selects a CHNLS_LIST-like subset (order preserved); emax
    keeps the pairs with E1 + E2 <= emax (cm^-1), also in that order -- the
    CHNLS_LIST of those pairs, not the energy-sorted EMAX= selection."""
    ld = 2 * max(jmax1, jmax2) + 1
    L1 = _levels(jmin1, jmax1, A1, B1, C1, ld)
    L2 = _levels(jmin2, jmax2, A2, B2, C2, ld)
    rows = []
    for a in L1:
        for b in L2:
            if keep is not None and not keep(a[0], a[1], a[2], b[0], b[1], b[2]):
                continue
            if emax is not None and a[3] + b[3] > emax:
                continue
            rows.append((a, b))
    return Basis(
        0,
        np.array([a[0] for a, b in rows], np.int32),
        np.array([a[3] + b[3] for a, b in rows]),
        np.array([a[1] for a, b in rows], np.int32),
        np.array([a[2] for a, b in rows], np.int32),
        np.array([b[0] for a, b in rows], np.int32),
        np.array([b[1] for a, b in rows], np.int32),
        np.array([b[2] for a, b in rows], np.int32),
        eig1=np.array([a[4] for a, b in rows]),
        eig2=np.array([b[4] for a, b in rows]),
    )


def channels_type0_list(rows, A1, B1, C1, A2=None, B2=None, C2=None):
    """Asym. top + asym. top channels from an explicit CHNLS_LIST of
    (j1, ka1, kc1, j2, ka2, kc2) rows, in the list's order."""
    A2, B2, C2 = (A1, B1, C1) if A2 is None else (A2, B2, C2)
    ld = 2 * max(max(r[0], r[3]) for r in rows) + 1
    out = []
    for j1, ka1, kc1, j2, ka2, kc2 in rows:
        E1, v1 = eigen_vec(A1, B1, C1, j1, ka1, kc1)
        E2, v2 = eigen_vec(A2, B2, C2, j2, ka2, kc2)
        a = np.zeros(ld)
        a[:2 * j1 + 1] = v1
        b = np.zeros(ld)
        b[:2 * j2 + 1] = v2
        out.append((j1, ka1, kc1, j2, ka2, kc2, E1 + E2, a, b))
    col = lambda i: np.array([r[i] for r in out], np.int32)
    return Basis(0, col(0), np.array([r[6] for r in out]), col(1), col(2), col(3), col(4), col(5),
                 eig1=np.array([r[7] for r in out]), eig2=np.array([r[8] for r in out]))


def inversion_sign(vec, J):
    """The sign of c_K c_{-K} for the first K = -J..0 where it is not zero
    (pot_comp.f:70-104): +1 when the level is even under K -> -K, -1 when
    odd."""
    for i in range(J + 1):
        p = vec[i] * vec[2 * J - i]
        if p != 0.0:
            return 1 if p > 0.0 else -1
        if abs(vec[i]) + abs(vec[2 * J - i]) != 0.0:
            raise ValueError("STOP: ASYMETRIC TOP DIAG WRONG")
    return 1


def kappa(ka, kc):
    """bikram_kappa (matrix_ini.f:9147-9180): (Ka - Kc) mod 2."""
    return abs(ka - kc) % 2


def ini_arrays(b: Basis, identical=False):
    """States (channel order, then j12, then m12 ascending) and the pair list
    st1 >= st2 with equal m12 (matrix_ini.f:255-630).  identical (type 0,
    IDENTICAL=YES): per channel the exchange parities p = +1, -1 outermost
    (only +1 when both partners are in the same level, where odd j12 states
    take -1), and pairs of equal m12 and equal exchange parity
    (matrix_ini.f:486-528, 556-630)."""
    if identical:
        return _ini_arrays_identical(b)
    j12, m12, ich, cidx = [], [], [], []
    for i in range(b.n_channels):
        if b.coll_type in (1, 4):
            j = int(b.j1_ch[i])
            for m in range(-j, j + 1):
                j12.append(j)
                m12.append(m)
                ich.append(i + 1)
                if m == 0:
                    cidx.append(len(j12))
        else:
            ja, jb = int(b.j1_ch[i]), int(b.j2_ch[i])
            first = None
            for js in range(abs(ja - jb), ja + jb + 1):
                for m in range(-js, js + 1):
                    j12.append(js)
                    m12.append(m)
                    ich.append(i + 1)
                    if m == 0 and js == abs(ja - jb):
                        first = len(j12)
            cidx.append(first)
    b.j12 = np.array(j12, np.int32)
    b.m12 = np.array(m12, np.int32)
    b.indx_chann = np.array(ich, np.int32)
    b.chann_indx = np.array(cidx, np.int32)
    b.Ej = np.asarray(b.E_ch)[b.indx_chann - 1]
    # pairs, vectorised per st1
    m = b.m12
    pairs = []
    for s1 in range(len(m)):
        s2 = np.nonzero(m[: s1 + 1] == m[s1])[0]
        pairs.append(np.stack([np.full(len(s2), s1 + 1), s2 + 1], 1))
    b.ind_mat = np.ascontiguousarray(np.concatenate(pairs).astype(np.int32))
    return b


def _ini_arrays_identical(b: Basis):
    if b.coll_type != 0:
        raise ValueError("ERROR: EXCHANGE SYMMETRY DEFINED FOR NON-IDENTICAL PARTICLES")
    j12, m12, ich, cidx, par, pbk = [], [], [], [], [], []
    for i in range(b.n_channels):
        ja, jb = int(b.j1_ch[i]), int(b.j2_ch[i])
        same = ja == jb and b.ka1_ch[i] == b.ka2_ch[i] and b.kc1_ch[i] == b.kc2_ch[i]
        p_inv = inversion_sign(b.eig1[i], ja) * inversion_sign(b.eig2[i], jb)
        k12 = (-1) ** (kappa(b.ka1_ch[i], b.kc1_ch[i]) + kappa(b.ka2_ch[i], b.kc2_ch[i]))
        first = None
        for p in range(1, (1 if same else 2) + 1):
            for js in range(abs(ja - jb), ja + jb + 1):
                for m in range(-js, js + 1):
                    j12.append(js)
                    m12.append(m)
                    ich.append(i + 1)
                    pbk.append((-1) ** js * p_inv * k12)
                    par.append(-1 if (same and js % 2 == 1) else (-1) ** (p - 1))
                    if m == 0 and js == abs(ja - jb) and p == 1:
                        first = len(j12)
        cidx.append(first)
    b.identical = True
    b.j12 = np.array(j12, np.int32)
    b.m12 = np.array(m12, np.int32)
    b.indx_chann = np.array(ich, np.int32)
    b.chann_indx = np.array(cidx, np.int32)
    b.parity_state = np.array(par, np.int32)
    b.parity_state_bk = np.array(pbk, np.int32)
    b.Ej = np.asarray(b.E_ch)[b.indx_chann - 1]
    m, pr = b.m12, b.parity_state
    pairs = []
    for s1 in range(len(m)):
        s2 = np.nonzero((m[: s1 + 1] == m[s1]) & (pr[: s1 + 1] == pr[s1]))[0]
        pairs.append(np.stack([np.full(len(s2), s1 + 1), s2 + 1], 1))
    b.ind_mat = np.ascontiguousarray(np.concatenate(pairs).astype(np.int32))
    return b


# ---------------------------------------------------------------------------
# grids
def gauss_legendre(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


@dataclass
class AngularGrid:
    """GRID_INTEGR state: nodes/weights on [-1,1] (or equidistant angles)
    and the affine maps of INI_POT_STORAGE."""
    n_beta: int
    xb: np.ndarray
    wb: np.ndarray
    xbm: float
    xbl: float
    n_gamma: int = 1
    xg: np.ndarray = None
    wg: np.ndarray = None
    xgm: float = 0.0
    xgl: float = 1.0
    n_alpha2: int = 1
    xaa: np.ndarray = None
    waa: np.ndarray = None
    xaam: float = 0.0
    xaal: float = 1.0
    n_beta2: int = 1
    xbb: np.ndarray = None
    wbb: np.ndarray = None
    xbbm: float = 0.0
    xbbl: float = 1.0
    n_gamma2: int = 1
    xgg: np.ndarray = None
    wgg: np.ndarray = None
    xggm: float = 0.0
    xggl: float = 1.0
    eq_gamma: int = 1
    ref_xz: int = 1

    @property
    def G(self):
        return self.n_beta * self.n_gamma * self.n_alpha2 * self.n_beta2 * self.n_gamma2


def _gl_0_pi(n):
    x, w = gauss_legendre(n)
    x2, x1 = PI, 0.0
    return x, w, 0.5 * (x2 + x1), 0.5 * (x2 - x1)


def _eq_0_2pi(n):
    """Equidistant angle grid (bikram_eq_grd_gamma, pot_comp.f:596-600)."""
    x1, xl = 0.0, PI
    i = np.arange(1, n + 1)
    return x1 + (i - 0.5) * 2.0 * xl / n, np.full(n, 2.0 / n), PI, PI


def grid_type1(n_beta):
    x, w, m, l = _gl_0_pi(n_beta)
    return AngularGrid(n_beta, x, w, m, l)


def grid_type4(n_beta, n_gamma, ref_xz=1):
    x, w, m, l = _gl_0_pi(n_beta)
    g, wg, gm, gl = _eq_0_2pi(n_gamma)
    return AngularGrid(n_beta, x, w, m, l, n_gamma, g, wg, gm, gl, ref_xz=ref_xz)


def grid_type0(n_alpha2, n_beta1, n_beta2, n_gamma1, n_gamma2):
    xb, wb, bm, bl = _gl_0_pi(n_beta1)
    xbb, wbb, bbm, bbl = _gl_0_pi(n_beta2)
    xg, wg, gm, gl = _eq_0_2pi(n_gamma1)
    xa, wa, am, al = _eq_0_2pi(n_alpha2)
    xgg, wgg, ggm, ggl = _eq_0_2pi(n_gamma2)
    return AngularGrid(n_beta1, xb, wb, bm, bl, n_gamma1, xg, wg, gm, gl,
                       n_alpha2, xa, wa, am, al, n_beta2, xbb, wbb, bbm, bbl,
                       n_gamma2, xgg, wgg, ggm, ggl)


def r_grid(n, rmin, rmax, equidistant=True):
    """R_COM (matrix_ini.f:954-1001): equidistant (the RGRID_EQDS default)
    or the three-section "optimized" grid."""
    R = np.zeros(n)
    if equidistant:
        d = (rmax - rmin) / (n - 1.0)
        for i in range(1, n + 1):
            R[i - 1] = rmin + d * (i - 1.0)
        return R
    n1 = n // 3
    n2 = n // 2 + 1
    R[0], R[n - 1] = rmin, rmax
    d = (rmax - rmin) / (n - 1.0)
    for i in range(1, n1 + 1):
        R[i - 1] = rmin + d * (i - 1.0) / 3.0
    for i in range(n1 + 1, n2 + 1):
        R[i - 1] = R[n1 - 1] + d * (i - n1)
    d = (rmax - R[n2 - 1]) / (n - n2)
    for i in range(n2, n + 1):
        R[i - 1] = R[n2 - 1] + d * (i - n2)
    return R


def nearest_r(R, r0):
    """minloc(|R - r0|), 1-based (matrix_ini.f:1969-1986)."""
    return int(np.argmin(np.abs(np.asarray(R) - r0))) + 1
