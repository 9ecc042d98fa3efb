"""The benchmark workloads: systems, grids and seeded synthetic potentials.

  diatom_atom      rigid diatom + atom (SYS_TYPE=1), B=1.93 cm^-1, mu=3.6 amu,
                   j=0..10 (11 channels, 121 states, 506 pairs), 64-point
                   Gauss-Legendre in theta, 200 R in 3..20 bohr,
                   V = sum_{l<=6} (A e^{-bR} - C/R^6) P_l(cos theta) + core.
  asym_top_atom    asymmetric top + atom (SYS_TYPE=4), A/B/C = 27.9/14.5/9.3,
                   j=0..10 (121 channels, 1771 states, 86 779 pairs),
                   beta 24 Gauss-Legendre x gamma 48 equidistant, 300 R in
                   3..20 bohr, V = real spherical harmonics l<=8 with the
                   same radial form, coefficients ~ 2^-l, + core.
  two_asym_tops    two water-like tops (SYS_TYPE=0), constants as above for
                   both, mu=9.0 amu, 3+3-site exp-6 + point charges,
                   16^5 angular grid, 128 R in 4..25 bohr.  The basis is
                   j1<=2, j2<=2 (81 channels, 1225 states) unless given.

Seeds are fixed per workload (numpy PCG64) so every run sees the same
potential.
"""
from dataclasses import dataclass

import numpy as np

from . import basis as B

SEEDS = {"diatom_atom": 20240501, "asym_top_atom": 20240502, "two_asym_tops": 20240503}

ROT_A, ROT_B, ROT_C = 27.9, 14.5, 9.3


# Isotropic repulsive core added to the Legendre and real-harmonic
# potentials: Aw exp(-bw (R - Rw)), 5000 cm^-1 at 3.5 bohr, below 0.25 cm^-1
# beyond 6 bohr.  The radial form A exp(-bR) - C/R^6 with A <= 1e5, b >= 1.5
# stays below 1.1e3 cm^-1 at R = 3 bohr and falls to -inf as R -> 0, so
# without a core the trajectories of the 200..2000 cm^-1 workloads with small
# impact parameters leave the R grid (R < R_1) into the spline's cubic
# extrapolation.  With it the turning point of a 2000 cm^-1 head-on
# collision lies near 3.7 bohr, inside the grid.
CORE = (5000.0, 4.0, 3.5)


def pes_legendre_params(seed, lmax=6, core=CORE):
    """[L, A_0..A_{L-1}, b_0.., C_0.., Aw, bw, Rw] for
    V = sum_l (A_l e^{-b_l R} - C_l / R^6) P_l(cos theta) + Aw e^{-bw (R - Rw)}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = lmax + 1
    A = rng.uniform(1e3, 1e5, L)
    b = rng.uniform(1.5, 2.5, L)
    C = rng.uniform(1e4, 1e6, L)
    return np.concatenate([[L], A, b, C, core]).astype(np.float64)


def pes_realsh_params(seed, lmax=8, core=CORE):
    """[lmax, (A_l, b_l, C_l, c_{l,-l..l}) for l = 0..lmax, Aw, bw, Rw] for
    V = sum_l (A_l e^{-b_l R} - C_l / R^6) sum_m c_lm S_lm(beta, gamma)
        + Aw e^{-bw (R - Rw)}, S_lm the real orthonormal spherical harmonics."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = [float(lmax)]
    for l in range(lmax + 1):
        A = rng.uniform(1e3, 1e5)
        b = rng.uniform(1.5, 2.5)
        C = rng.uniform(1e4, 1e6)
        c = rng.uniform(-1.0, 1.0, 2 * l + 1) * 2.0 ** (-l)
        out += [A, b, C, *c]
    out += list(core)
    return np.array(out, np.float64)


def water_sites(rng):
    """O and two H in a body frame with the centre of mass at the origin
    (bohr); charges neutral, |q_H| in [0.3, 0.4] e; exp-6 eps in 10..200
    cm^-1 and r0 in 2.5..4 bohr."""
    roh, ang = 1.8088, np.deg2rad(104.52)
    mO, mH = 15.995, 1.008
    h1 = np.array([roh * np.sin(ang / 2), 0.0, roh * np.cos(ang / 2)])
    h2 = np.array([-roh * np.sin(ang / 2), 0.0, roh * np.cos(ang / 2)])
    o = np.zeros(3)
    com = (mO * o + mH * h1 + mH * h2) / (mO + 2 * mH)
    qh = rng.uniform(0.3, 0.4)
    sites = []
    for p, q in ((o - com, -2 * qh), (h1 - com, qh), (h2 - com, qh)):
        sites.append([p[0], p[1], p[2], q, rng.uniform(10.0, 200.0), rng.uniform(2.5, 4.0)])
    return sites


def pes_sites_params(seed, alpha=14.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    s1 = water_sites(rng)
    s2 = water_sites(rng)
    return np.array([alpha, 3, 3, *np.ravel(s1), *np.ravel(s2)], np.float64)


@dataclass
class Workload:
    name: str
    basis: B.Basis
    grid: B.AngularGrid
    R: np.ndarray
    pes_kind: int
    pes_par: np.ndarray
    mass_amu: float
    conv_unit_e: float = B.CM_TO_HARTREE
    mij_cutoff_cm: float = 1e-12          # MIJ_CUTOFF default (iotest.f:3787)
    cutoff_r1: float = 6.0                # CUTOFF_R1 default (iotest.f:3788)
    cutoff_r2: float = 8.0                # CUTOFF_R2 default (iotest.f:3789)

    @property
    def mij_zero_cut(self):
        return self.mij_cutoff_cm * B.CM_TO_HARTREE

    @property
    def ir_cut(self):
        return B.nearest_r(self.R, self.cutoff_r1), B.nearest_r(self.R, self.cutoff_r2)

    def summary(self):
        b = self.basis
        return {"workload": self.name, "sys_type": b.coll_type, "channels": b.n_channels,
                "states": b.n_states, "pairs": int(b.total_size), "grid_points": int(self.grid.G),
                "r_points": int(len(self.R))}


def diatom_atom(jmax=10, n_beta=64, n_r=200, seed=None):
    b = B.ini_arrays(B.channels_type1(0, jmax, 1.93))
    g = B.grid_type1(n_beta)
    R = B.r_grid(n_r, 3.0, 20.0)
    par = pes_legendre_params(SEEDS["diatom_atom"] if seed is None else seed)
    return Workload("diatom_atom", b, g, R, 0, par, 3.6)


def asym_top_atom(jmax=10, n_beta=24, n_gamma=48, n_r=300, seed=None, ref_xz=1):
    b = B.ini_arrays(B.channels_type4(0, jmax, ROT_A, ROT_B, ROT_C))
    g = B.grid_type4(n_beta, n_gamma, ref_xz=ref_xz)
    R = B.r_grid(n_r, 3.0, 20.0)
    par = pes_realsh_params(SEEDS["asym_top_atom"] if seed is None else seed)
    return Workload("asym_top_atom", b, g, R, 1, par, 3.27)


# The two-tops basis of the benchmark: the CHNLS_LIST of every pair level
# with j1, j2 <= 4 and E1 + E2 <= 236.5 cm^-1, molecule 1 outermost (123
# channels, 2043 states, 282 179 pairs; m-blocks of up to 369 states).  The
# literal "j1, j2 <= 4 with all m and tau" is 27 225 states and 32.8 M
# pairs; an energy cut over whole pair levels reaches 2001 or 2043 states
# around the stated 2025, and the cut above the 236.41 cm^-1 levels is taken.
C3_JMAX, C3_EMAX = 4, 236.5


def two_asym_tops(jmax1=2, jmax2=2, n_ang=16, n_r=128, seed=None, keep=None,
                  n_ang_each=None, emax=None):
    b = B.ini_arrays(B.channels_type0(0, jmax1, 0, jmax2, ROT_A, ROT_B, ROT_C,
                                      ROT_A, ROT_B, ROT_C, keep=keep, emax=emax))
    na = n_ang_each or (n_ang,) * 5
    g = B.grid_type0(*na)
    R = B.r_grid(n_r, 4.0, 25.0)
    par = pes_sites_params(SEEDS["two_asym_tops"] if seed is None else seed)
    return Workload("two_asym_tops", b, g, R, 2, par, 9.0)


def two_asym_tops_c3(n_ang=16, n_r=128, seed=None):
    """The benchmark's two-tops system: C3_JMAX, C3_EMAX pair levels on the
    16^5 grid with 128 R from 4 to 25 bohr."""
    return two_asym_tops(C3_JMAX, C3_JMAX, n_ang=n_ang, n_r=n_r, seed=seed, emax=C3_EMAX)


def at_batch(n_traj=4096, b_max=18.0, e_min=200.0, e_max=2000.0, seed=20240504):
    """The AT-MQCT propagation workload: trajectories of the asymmetric
    top + atom system with impact parameters uniform in [0, b_max] bohr and
    collision energies uniform in [e_min, e_max] cm^-1 (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = rng.uniform(0.0, b_max, n_traj)
    E = rng.uniform(e_min, e_max, n_traj)
    return b, E


def two_tops_at_batch(n_traj=1024, b_max=15.0, e_min=200.0, e_max=1500.0, seed=20240505):
    """AT-MQCT trajectories of the two-asymmetric-tops system: impact
    parameters uniform in [0, b_max] bohr and collision energies uniform in
    [e_min, e_max] cm^-1 (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(0.0, b_max, n_traj), rng.uniform(e_min, e_max, n_traj)

