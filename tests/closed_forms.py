"""Closed forms the parity tests pin the matrix elements with (test helpers,
not tests).

Asymmetric top + atom (SYS_TYPE 4).  The reference's wavefunction is
psi_{J M tau}(alpha, beta, gamma) = N_J sum_K c_K Djkm(J, K, M), where
Djkm(j, k, m) = d^j_{m k}(beta) exp(i (k gamma + m alpha))
(pot_comp.f:256-309, basis_wave_functions.f:83-116), i.e. the complex
conjugate of the standard D^J_{M K}(alpha, beta, gamma) =
exp(-i M alpha) d^J_{M K}(beta) exp(-i K gamma), with N_J =
sqrt((2J+1)/8 pi^2).  For a potential that is a sum of
D^l_{0 q}(alpha, beta, gamma) = d^l_{0 q}(beta) exp(-i q gamma) (a function
of beta and gamma only), the integral over the three Euler angles is

  int psi_1 psi_2^* D^l_{0q} dOmega
     = sqrt((2J1+1)(2J2+1)) sum_{K1,K2} c1_K1 c2_K2 (-1)^(M-K1)
       (J1 J2 l; -M M 0) (J1 J2 l; -K1 K2 q)

(D^{J*}_{M K} = (-1)^(M-K) D^J_{-M -K}, and the integral of three D
functions is 8 pi^2 times two 3j symbols).  The reference integrates beta
and gamma on the grid and alpha analytically (2 pi), and takes the real
part (REF_XZ=YES) or real + imaginary part (REF_XZ=NO, pot_comp.f:2544-2553)
of that integral.  The integrand has trigonometric terms in gamma of order
<= J1 + J2 + l, which an equidistant gamma grid of m points integrates
exactly when J1 + J2 + l < m; in beta the reference's Gauss-Legendre nodes
lie in beta itself (sin(beta) is a factor of the integrand), so that sum
converges exponentially with the number of nodes rather than being exact.
"""
import functools
import math
from fractions import Fraction

import numpy as np


@functools.lru_cache(maxsize=None)
def exact_3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol by the Racah sum in exact rationals (integer j),
    rounded once at the end."""
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    pre = tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    s = Fraction(0)
    for k in range(0, j1 + j2 + j3 + 1):
        a = [k, j1 + j2 - j3 - k, j1 - m1 - k, j2 + m2 - k, j3 - j2 + m1 + k, j3 - j1 - m2 + k]
        if min(a) < 0:
            continue
        den = 1
        for x in a:
            den *= f(x)
        s += Fraction((-1) ** k, den)
    if s == 0:
        return 0.0
    sign = (-1) ** (j1 - j2 - m3) * (1 if s > 0 else -1)
    return sign * math.sqrt(s * s * pre)


def wigner_d(j, mp, m, b):
    """d^j_{m' m}(b) by the explicit Wigner sum (standard convention)."""
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


def d_potential_coeffs(rng, lmax):
    """Random complex weights w[l][q] (q = -l..l) of
    V = Re sum_{l,q} w_lq D^l_{0q}(beta, gamma)."""
    return [rng.standard_normal(2 * l + 1) + 1j * rng.standard_normal(2 * l + 1) for l in range(lmax + 1)]


def d_potential_grid(w, beta, gamma):
    """V[beta][gamma] of the potential with weights w."""
    V = np.zeros((len(beta), len(gamma)), dtype=complex)
    for l, wl in enumerate(w):
        for iq, q in enumerate(range(-l, l + 1)):
            db = np.array([wigner_d(l, 0, q, b) for b in beta])
            V += wl[iq] * db[:, None] * np.exp(-1j * q * np.asarray(gamma))[None, :]
    return V.real


def type4_element(w, J1, c1, J2, c2, M, ref_xz=1):
    """int psi_1 psi_2^* V over alpha, beta, gamma for V = Re sum w_lq
    D^l_{0q}; c1, c2: coefficients c_K, K = -J..J.  Returns Re(I)
    (REF_XZ=YES) or Re(I) + Im(I) (REF_XZ=NO)."""
    def T(l, q):
        s = 0.0
        for i1, K1 in enumerate(range(-J1, J1 + 1)):
            K2 = K1 - q
            if abs(K2) > J2 or c1[i1] == 0.0:
                continue
            s += c1[i1] * c2[K2 + J2] * (-1) ** (M - K1) * exact_3j(J1, J2, l, -K1, K2, q)
        return s * math.sqrt((2 * J1 + 1) * (2 * J2 + 1)) * exact_3j(J1, J2, l, -M, M, 0)

    I = 0.0 + 0.0j
    for l, wl in enumerate(w):
        for iq, q in enumerate(range(-l, l + 1)):
            # Re(w D) = (w D + w^* D^*)/2 and D^{l*}_{0q} = (-1)^q D^l_{0,-q}
            I += 0.5 * (wl[iq] * T(l, q) + np.conj(wl[iq]) * (-1) ** q * T(l, -q))
    return I.real if ref_xz else I.real + I.imag


def h2ohe_reference_system():
    """The reference's SYS_TYPE 4 example EXAMPLES_IO/SYS_TYPE_4/
    H2O_He_external_rot_wfn as data: its 9 channels with the explicit
    eigenvectors of USER_DEFINED_BASIS.DAT (LEVELS_FILE=YES), the 10 x 20
    beta x gamma grid and R = 6, 8 bohr of its deck, the V grid of its own
    PES (compiled from source; tests/golden/make_h2ohe_vgrid.py) and the
    Mij of its MTRX.DAT (Hartree, 12 significant digits).  Returns
    (basis, grid, R, V[R][beta][gamma] in cm^-1, mij[k][iR])."""
    import json
    import os
    from mqct_amd import basis as B
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    fx = json.load(open(os.path.join(here, "fixtures.json")))["h2o_he"]
    vg = json.load(open(os.path.join(here, "h2ohe_vgrid.json")))
    ch = np.array(fx["channels"])
    j1 = ch[:, 0].astype(np.int32)
    eig = np.zeros((len(ch), 2 * int(j1.max()) + 1))
    for c, v in enumerate(fx["eigvecs"]):
        eig[c, :len(v)] = v
    b = B.ini_arrays(B.Basis(4, j1, ch[:, 3], ch[:, 1].astype(np.int32), ch[:, 2].astype(np.int32), eig1=eig))
    st = np.array(fx["states"])
    assert np.array_equal(b.j12, st[:, 2]) and np.array_equal(b.m12, st[:, 3])
    assert np.array_equal(b.ind_mat, np.array(fx["pairs"])[:, 1:3])
    g = B.grid_type4(vg["grid"]["n_beta"], vg["grid"]["n_gamma"])
    return b, g, np.array(vg["R"]), np.array(vg["V"]), np.array(vg["mij"])


def h2o_h2o_identical_reference_system():
    """The reference's SYS_TYPE 0 example EXAMPLES_IO/SYS_TYPE_0/
    H2O_H2O_Identical as data (copied under tests/golden/expansion): the
    input's CHNLS_LIST, rotational constants and 152 IDENTICAL_PES terms,
    its PES_EXPAN_TERMS.DAT (2 R points, kcal/mol) and the states, pairs and
    Mij (Hartree, 12 digits) of its MTRX.DAT.  Returns a dict."""
    import os
    import re
    from mqct_amd import basis as B
    from mqct_amd import expansion as X
    d = os.path.join(os.path.dirname(__file__), "golden", "expansion")
    inp = open(os.path.join(d, "h2o_h2o_identical.inp")).read()
    num = lambda k: float(re.search(k + r"\s*=\s*([-0-9.]+)", inp).group(1))
    A, Bc, C = num("A1"), num("B1"), num("C1")
    i0 = inp.index("CHNLS_LIST=")
    rows = re.findall(r"^\s*(\d+),(\d+),(\d+),(\d+),(\d+),(\d+),\s*$", inp[i0:inp.index("INIT_CHNL")], re.M)
    rows = [tuple(int(v) for v in r) for r in rows]
    i1 = inp.index("TERMS =")
    terms = X.parse_terms_list(inp[inp.index("\n", i1) + 1:inp.index("$END", i1)])
    R, coeffs = X.read_pes_expan_terms(os.path.join(d, "h2o_h2o_identical_PES_EXPAN_TERMS.DAT"), len(terms))
    lines = open(os.path.join(d, "h2o_h2o_identical_MTRX.DAT")).read().splitlines()
    ns = int(lines[2].split("=")[1])
    T = int(lines[3].split("=")[1])
    st = np.array([[int(x) for x in ln.split()] for ln in lines[6:6 + ns]])
    pr = np.array([[int(x) for x in ln.split()] for ln in lines[7 + ns:7 + ns + T]])
    el = lines[7 + ns + T + 1 + len(R) + 1:]
    mij = np.zeros((T, len(R)))
    for ln in el:
        k, ir, v = ln.split()
        mij[int(k) - 1, int(ir) - 1] = float(v)
    return {"rows": rows, "A": A, "B": Bc, "C": C, "terms": terms, "R": R, "coeffs": coeffs,
            "states": st, "pairs": pr[:, 1:], "mij": mij, "conv_unit_e": 1.0 / 627.509469}
