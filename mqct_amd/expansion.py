"""PES expansion through the HIP library: the host side of CALC_EXPANSION
(pot_comp.f:3898-4693), CALC_EXPANSION_TERMS (pot_comp.f:4695-6274),
READ_EXPANSION_TERMS (pot_comp.f:6276-6395) and MATRIX_MAKE's expansion
branch (matrix_ini.f:1950-1956 -> EXPANSION_MATRIX_ELEMENT,
pot_comp.f:6465-8751).

    with Device() as dev:
        mb = MatrixBuilder(dev, workload)               # basis, grid, V grid
        ex = Expansion(dev, workload, terms_type4(10, 6, du1=2))
        coeffs = ex.project()                           # [n_terms][n_r]
        mat_el, mat_el_der, skipped = ex.build()        # from the projection
        write_pes_expan_terms("PES_EXPAN_TERMS.DAT", 4, R, ex.a_top, coeffs)

Term rows are A_TOP(1:5, t): type 1 (L,), type 4 (L, NJU), type 0
(L1, NJU1, L2, NJU2, L).  Coefficients follow the reference's
expansion_terms(n_r, n_terms): coeffs[t, ir].
"""
import numpy as np

from . import _lib as L


# ---- term lists as CALC_EXPANSION generates them ---------------------------
def terms_type1(l_max, dlamda1=1):
    """SYS_TYPE 1: L = 0..L_MAX step EQU_SYM1 (pot_comp.f:3926-3936)."""
    return [(l,) for l in range(0, l_max + 1, dlamda1)]


def terms_type4(l_max, nju_max=-1, du1=1, dlamda1=1):
    """SYS_TYPE 4: L = 0..L_MAX step EQU_SYM1, NJU = 0..min(M_MAX, L) step
    AXL_SYM1 (pot_comp.f:4057-4080); M_MAX < 0 means L_MAX."""
    if nju_max < 0:
        nju_max = l_max
    return [(l, n) for l in range(0, l_max + 1, dlamda1) for n in range(0, min(nju_max, l) + 1, du1)]


def terms_type0(l1_max, l2_max, l_max=-1, nju1_max=-1, nju2_max=-1, l1_min=0, l2_min=0, du1=1, du2=1,
                dlamda1=1, dlamda2=1, odd_l1l2l=False):
    """SYS_TYPE 0, distinguishable partners (pot_comp.f:4470-4500): L1, NJU1
    >= 0, L2, NJU2 in -min(L2, NJU2_MAX)..min(L2, NJU2_MAX) (NJU2 >= 0 when
    NJU1 = 0), L from |L1-L2| to min(L1+L2, L_MAX); L1+L2+L even unless
    ODD_L1L2L."""
    if nju1_max < 0:
        nju1_max = l1_max
    if nju2_max < 0:
        nju2_max = l2_max
    if l_max < 0:
        l_max = l1_max + l2_max
    out = []
    for l1 in range(l1_min, l1_max + 1, dlamda1):
        for n1 in range(0, min(l1, nju1_max) + 1, du1):
            for l2 in range(l2_min, l2_max + 1, dlamda2):
                m2 = min(l2, nju2_max)
                for n2 in range(-m2, m2 + 1, du2):
                    for l in range(abs(l1 - l2), min(l1 + l2, l_max) + 1):
                        if n1 == 0 and n2 < 0:
                            continue
                        if not odd_l1l2l and (l1 + l2 + l) % 2 != 0:
                            continue
                        out.append((l1, n1, l2, n2, l))
    return out


def format_terms_list(terms):
    """EXPANSION_TERMS_LIST.DAT text: one term per line, each value followed
    by a comma except the last of a two-value type-4 row (the formats
    '(1(i0,a))', '(2(i0,a))' and '(5(i0,a))' of pot_comp.f:3930, 4071,
    4494 with the values the reference passes them)."""
    lines = []
    for t in terms:
        if len(t) == 2:
            lines.append("%d,%d" % t)
        else:
            lines.append("".join("%d," % v for v in t))
    return "\n".join(lines) + "\n"


def parse_terms_list(text):
    """TERMS= / EXPANSION_TERMS_LIST.DAT rows back to tuples."""
    out = []
    for line in text.splitlines():
        vals = [v for v in line.replace(" ", "").split(",") if v != ""]
        if vals:
            out.append(tuple(int(v) for v in vals))
    return out


def a_top_rows(terms):
    """A_TOP(1:5, t) rows, zero-padded."""
    a = np.zeros((len(terms), 5), np.int32)
    for i, t in enumerate(terms):
        a[i, :len(t)] = t
    return a


# ---- PES_EXPAN_TERMS.DAT -----------------------------------------------------
def _e19_12(x):
    """Fortran E19.12: [sign]0.dddddddddddd followed by E+xx (or, for a
    three-digit exponent, the sign and three digits without the E)."""
    if x == 0.0:
        return " 0.000000000000E+00"
    m, e = ("%.11e" % abs(x)).split("e")
    digits = m.replace(".", "")
    ex = int(e) + 1
    if abs(ex) <= 99:
        tail = "E" + ("+" if ex >= 0 else "-") + "%02d" % abs(ex)
    else:
        tail = ("+" if ex >= 0 else "-") + "%03d" % abs(ex)
    return (("-" if x < 0 else " ") + "0." + digits + tail).rjust(19)


def _f_real(field):
    """A Fortran real field (E or D exponent letter, or a bare signed
    three-digit exponent)."""
    f = field.strip().replace("D", "E").replace("d", "E")
    if "E" not in f and len(f) > 4 and f[-4] in "+-" and f[-5].isdigit():
        f = f[:-4] + "E" + f[-4:]
    return float(f)


def write_pes_expan_terms(path, coll_type, R, terms, coeffs):
    """PES_EXPAN_TERMS.DAT as CALC_EXPANSION_TERMS writes it
    (pot_comp.f:5962-6010 for types 1 and 4, 6230-6264 for type 0): a header
    of term labels, then per R point '(f6.2,1x,e19.12,3x)' and
    '(e19.12,3x)' per further term."""
    coeffs = np.asarray(coeffs, dtype=np.float64)
    with open(path, "w") as f:
        labels = ["v_" + "_".join(str(v) for v in t[:{1: 1, 4: 2, 0: 5}[coll_type]]) for t in terms]
        f.write("%-6s %-19s   " % ("R_Bohr", labels[0]) + "".join("%-19s   " % s for s in labels[1:]) + "\n")
        for ir, r in enumerate(R):
            f.write("%6.2f " % r)
            f.write("".join(_e19_12(coeffs[t, ir]) + "   " for t in range(len(terms))))
            f.write("\n")


def read_pes_expan_terms(path, n_terms, n_r=None):
    """READ_EXPANSION_TERMS (pot_comp.f:6361-6374): skip the header, then per
    R point the fixed fields f6.2, 1x, e19.12, 3x, (e19.12, 3x) ...
    Returns (R, coeffs[n_terms][n_r]); R keeps the file's two decimals as
    the reference's read does."""
    with open(path) as f:
        lines = f.read().splitlines()[1:]
    lines = [ln for ln in lines if ln.strip()]
    if n_r is not None:
        lines = lines[:n_r]
    R = np.zeros(len(lines))
    C = np.zeros((n_terms, len(lines)))
    for ir, ln in enumerate(lines):
        R[ir] = float(ln[0:6])
        pos = 7
        for t in range(n_terms):
            C[t, ir] = _f_real(ln[pos:pos + 19])
            pos += 22
    return R, C


# ---- the library path --------------------------------------------------------
class Expansion:
    """The expansion path of one handle whose basis (and, for project(), grid
    and V grid) a MatrixBuilder has set."""

    def __init__(self, dev, wl, terms, identical_pes=False):
        """identical_pes (IDENTICAL_PES=YES): the terms are those of an
        exchange-symmetric PES (L2 <= L1); implied for a basis of identical
        partners, which MatrixBuilder has declared to the handle."""
        self.dev, self.lib, self.h = dev, dev.lib, dev.h
        self.wl = wl
        self.terms = [tuple(t) for t in terms]
        self.a_top = a_top_rows(self.terms)
        if identical_pes and not getattr(wl.basis, "identical", False):
            L.check(self.lib.mqct_hip_set_identical(self.h, 1, None, None, None), "set_identical")
        L.check(self.lib.mqct_hip_set_expansion_terms(self.h, len(self.terms), L.iptr(self.a_top)),
                "set_expansion_terms")

    def project(self):
        """CALC_EXPANSION_TERMS on the handle's V grid: [n_terms][n_r]."""
        C = np.zeros((len(self.terms), len(self.wl.R)))
        L.check(self.lib.mqct_hip_expansion_project(self.h, L.dptr(C)), "expansion_project")
        return C

    def build(self, coeffs=None, R=None, copy_out=True):
        """MATRIX_MAKE with the expansion: from coeffs on R (as read from
        PES_EXPAN_TERMS.DAT), or from the last projection when coeffs is
        None.  Returns (mat_el, mat_el_der, skipped) like MatrixBuilder."""
        b = self.wl.basis
        ind = L.as_i32(b.ind_mat)
        ir1, ir2 = self.wl.ir_cut
        if coeffs is not None:
            C = L.as_f64(coeffs)
            Rr = L.as_f64(self.wl.R if R is None else R)
            n_r, pR, pC = len(Rr), L.dptr(Rr), L.dptr(C)
        else:
            n_r, pR, pC = len(self.wl.R), None, None
        T = b.total_size
        if not copy_out:
            L.check(self.lib.mqct_hip_expansion_build(self.h, n_r, pR, pC, T, L.iptr(ind), self.wl.conv_unit_e,
                                                      self.wl.mij_zero_cut, ir1, ir2, None, None, None),
                    "expansion_build")
            return None
        M, D = np.zeros((T, n_r)), np.zeros((T, n_r))
        S = np.zeros(T, np.uint8)
        L.check(self.lib.mqct_hip_expansion_build(self.h, n_r, pR, pC, T, L.iptr(ind), self.wl.conv_unit_e,
                                                  self.wl.mij_zero_cut, ir1, ir2, L.dptr(M), L.dptr(D),
                                                  S.ctypes.data_as(L.c_u8_p)), "expansion_build")
        return M, D, S

    def timing(self):
        t = np.zeros(4)
        L.check(self.lib.mqct_hip_expansion_timing(self.h, L.dptr(t)), "expansion_timing")
        return {"project_terms_ms": t[0], "project_gemm_ms": t[1], "angular_ms": t[2], "gemm_ms": t[3]}
