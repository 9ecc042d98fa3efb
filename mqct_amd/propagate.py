"""Propagation through the HIP library: the host-side mirror of PROPAGATE /
DOTRAJECT (propagator.f) for a batch of trajectories, and the cross-section
assembly of INELAST_CALC (propagator.f:4752-4983, 1198-1276).

    prop = Propagator(dev, basis, R, mat_el, mat_el_der, mass_amu=3.6)
    trajs = prop.l_grid_trajectories(U_cm=500, b_max=15, s_st=1, time_step=10)
    q, stats = prop.run(MQCT_PROP_CC_RK4, trajs)
    P = prop.opacities(q)                      # per trajectory, per channel
    sigma = cross_sections(...)
"""
import ctypes

import numpy as np

from . import _lib as L
from .basis import A_BOHR, AMUTOAU, CM_TO_HARTREE

ENER_REF = 300.0  # propagator.f:46


def traj_array(rows):
    arr = (L.TrajT * len(rows))()
    for i, r in enumerate(rows):
        for k, v in r.items():
            setattr(arr[i], k, v)
    return arr


def l_grid(mass_au, U_cm, b_max, jmin, jmax, dl=1):
    """L values of PROPAGATE (propagator.f:467-494) and TRAJ_ORB
    (L = L_min + (id-1)*DL): k = sqrt(2 mu U), J_max = int(b_max k) - j_max,
    L from the smallest |J - j| to J_max + j_max.  Returns (L values,
    (J_min, J_max), k)."""
    k = np.sqrt(mass_au * 2.0 * U_cm * CM_TO_HARTREE)
    J_max = int(b_max * k) - jmax
    J_min = 0
    L_max = J_max + jmax
    L_min = min(abs(J - jj) for J in range(J_min, J_max + 1) for jj in range(jmin, jmax + 1))
    n = (L_max - L_min) // dl + 1
    return np.array([L_min + i * dl for i in range(n)]), (J_min, J_max), k


class Propagator:
    def __init__(self, dev, basis, R, mat_el=None, mat_el_der=None, mass_amu=1.0, theta_on=0,
                 from_build=False):
        """Make the spline coefficients resident: either host arrays
        (READ_MTRX path) or the device result of the last matrix build."""
        self.dev, self.lib, self.h = dev, dev.lib, dev.h
        self.b = basis
        self.R = L.as_f64(R)
        self.mass = mass_amu * AMUTOAU
        self.N = basis.n_states
        Ej = L.as_f64(np.asarray(basis.Ej) * CM_TO_HARTREE)
        self.Ej = Ej
        if from_build:
            L.check(self.lib.mqct_hip_use_built_matrix(self.h, L.dptr(Ej), self.mass, theta_on),
                    "use_built_matrix")
        else:
            M, D = L.as_f64(mat_el), L.as_f64(mat_el_der)
            L.check(self.lib.mqct_hip_load_matrix(self.h, self.N, L.dptr(Ej), L.iptr(L.as_i32(basis.j12)),
                                                  L.iptr(L.as_i32(basis.m12)), basis.total_size,
                                                  L.iptr(L.as_i32(basis.ind_mat)), len(self.R), L.dptr(self.R),
                                                  L.dptr(M), L.dptr(D), self.mass, theta_on), "load_matrix")

    # -- trajectory sets -----------------------------------------------------
    def l_grid(self, U_cm, b_max, chann_ini, dl=1):
        """L values for initial channel chann_ini (1-based); see l_grid()."""
        j = int(self.b.j12[self.b.chann_indx[chann_ini - 1] - 1]) if self.b.coll_type != 0 else None
        jmin = jmax = j if j is not None else None
        if self.b.coll_type == 0:
            states = np.nonzero(self.b.indx_chann == chann_ini)[0]
            jmin, jmax = int(self.b.j12[states].min()), int(self.b.j12[states].max())
        return l_grid(self.mass, U_cm, b_max, jmin, jmax, dl)

    def trajectories(self, l_values, U_cm, s_st, time_step=10.0, eps=1e-8, hmin=1e-6,
                     R_start=None, numb_oscl_prds=1, numb_orb_prds=1, at_paths=None, max_steps=200000):
        E = U_cm * CM_TO_HARTREE
        dt = time_step / np.sqrt(U_cm / ENER_REF)  # propagator.f:2436
        R_start = float(self.R[-1]) if R_start is None else R_start
        rows = []
        for i, l in enumerate(l_values):
            rows.append(dict(s_st=int(s_st), at_path=int(at_paths[i]) if at_paths is not None else 0,
                             l_real=float(l), E_coll=float(E if np.isscalar(E) else E[i]), dt=float(dt),
                             eps=float(eps), hmin=float(hmin), R_start=float(R_start), time_lim=-1.0,
                             numb_oscl_prds=numb_oscl_prds, numb_orb_prds=numb_orb_prds,
                             max_steps=max_steps))
        return rows

    # -- runs ---------------------------------------------------------------
    def run(self, mode, rows, batched=False):
        n = len(rows)
        arr = traj_array(rows)
        q = np.zeros((n, 2 * self.N + 8))
        st = (L.TrajStatT * n)()
        L.check(self.lib.mqct_hip_propagate(self.h, mode, n, arr, int(batched), L.dptr(q), st), "propagate")
        return q, stats_dict(st)

    def run_record(self, rows, rec_cap):
        n = len(rows)
        arr = traj_array(rows)
        q = np.zeros((n, 2 * self.N + 8))
        rec = np.zeros((n, rec_cap, 8))
        rn = np.zeros(n, np.int32)
        st = (L.TrajStatT * n)()
        L.check(self.lib.mqct_hip_propagate_record(self.h, n, arr, rec_cap, L.dptr(rec), L.iptr(rn), L.dptr(q),
                                                   st), "propagate_record")
        return q, stats_dict(st), [rec[i, :rn[i]].copy() for i in range(n)]

    def load_paths(self, paths):
        off = np.zeros(len(paths) + 1, np.int64)
        off[1:] = np.cumsum([len(p) for p in paths])
        recs = L.as_f64(np.concatenate(paths))
        L.check(self.lib.mqct_hip_load_paths(self.h, len(paths), off.ctypes.data_as(L.c_i64_p), L.dptr(recs)),
                "load_paths")

    def last_ms(self):
        t = ctypes.c_double()
        L.check(self.lib.mqct_hip_propagate_timing(self.h, ctypes.byref(t)), "propagate_timing")
        return t.value

    # -- results ------------------------------------------------------------
    def opacities(self, q):
        """sum |a|^2 over the states of each channel (propagator.f:4364-4383)."""
        N = self.N
        p = q[:, :N] ** 2 + q[:, N:2 * N] ** 2
        P = np.zeros((q.shape[0], self.b.n_channels))
        np.add.at(P.T, self.b.indx_chann - 1, p.T)
        return P


def stats_dict(st):
    return {"t_end": np.array([s.t_end for s in st]), "n_steps": np.array([s.n_steps for s in st]),
            "n_rhs": np.array([s.n_rhs for s in st]), "fail": np.array([s.fail for s in st]),
            "norm_err": np.array([s.norm_err for s in st])}


def inelastic_sigma(P, l_values, J_range, j_ini, k):
    """INELAST_CALC (propagator.f:4778-4833) without identical partners:
    sigma_n = sum_J sum_{l=|J-j|}^{J+j} P_n(l) (2J+1) pi / (k^2 (2j+1)), a.u.
    P: [trajectory(l index), channel]."""
    lmap = {int(l): i for i, l in enumerate(l_values)}
    sig = np.zeros(P.shape[1])
    for J in range(J_range[0], J_range[1] + 1):
        for l in range(abs(J - j_ini), J + j_ini + 1):
            if l not in lmap:
                continue
            sig += P[lmap[l]] * (2.0 * J + 1.0) / k ** 2 / (2 * j_ini + 1) * np.pi
    return sig


def inelastic_weights(l_values, J_range, j_ini, k, per_J=False):
    """The INELAST_CALC factors (propagator.f:4778-4833) as weights per
    trajectory: sigma_n = sum_t w[t] P_n(t) with w[t] = sum over the J whose
    l range |J - j| .. J + j holds l_t of (2J+1) pi / (k^2 (2j+1)).
    per_J: one row per J of J_range (the J-resolved table) instead.
    Returns [groups][trajectories]."""
    lmap = {int(l): i for i, l in enumerate(l_values)}
    Js = list(range(J_range[0], J_range[1] + 1))
    W = np.zeros((len(Js) if per_J else 1, len(l_values)))
    for gi, J in enumerate(Js):
        for l in range(abs(J - j_ini), J + j_ini + 1):
            if l in lmap:
                W[gi if per_J else 0, lmap[l]] += (2.0 * J + 1.0) / k ** 2 / (2 * j_ini + 1) * np.pi
    return W


def gpu_cross_sections(dev, q, basis, weights, opacity=False):
    """sigma[g][n] = sum_t weights[g][t] P_n(t) from the final amplitudes q
    (rows as Propagator.run returns them) on the GPU
    (mqct_hip_cross_sections); bohr^2.  With opacity=True also returns
    P[t][n]."""
    q = L.as_f64(q)
    W = L.as_f64(np.atleast_2d(weights))
    T, ldq = q.shape
    C = basis.n_channels
    sig = np.zeros((W.shape[0], C))
    P = np.zeros((T, C)) if opacity else None
    L.check(dev.lib.mqct_hip_cross_sections(dev.h, T, basis.n_states, ldq, L.dptr(q), C,
                                            L.iptr(L.as_i32(basis.indx_chann)), W.shape[0], L.dptr(W),
                                            L.dptr(sig), L.dptr(P)), "cross_sections")
    return (sig, P) if opacity else sig


def billing_scale(E_ch_cm, chann_ini, U_cm):
    """a0^2 U / E_bill with E_bill = U + dE/2 + dE^2/(16U) (propagator.f:276-287, 1208)."""
    dE = np.asarray(E_ch_cm) - E_ch_cm[chann_ini - 1]
    E_bill = U_cm + dE / 2.0 + dE ** 2 / 16.0 / U_cm
    return A_BOHR ** 2 * U_cm / E_bill


def m_average_weight(m, jmin_ind, jmax_ind):
    """propagator.f:1268-1274: 2/(1+delta_m0) / ((jmax+jmin+1)(jmax-jmin+1))."""
    return 2.0 / (1.0 + (1.0 if m == 0 else 0.0)) / ((jmax_ind + jmin_ind + 1.0) * (jmax_ind - jmin_ind + 1.0))


def write_at_paths(directory, paths, j12, m12, l_values, parity=0):
    """Writes every path to directory/AT_<j12>_<m12>_<L>_<parity> and
    n_AT_<...> in the reference's record layout (mqct_hip_write_at_path)."""
    lib = L.load()
    for p, l in zip(paths, l_values):
        rec = L.as_f64(p)
        L.check(lib.mqct_hip_write_at_path(str(directory).encode(), int(j12), int(m12), int(l), int(parity),
                                           len(rec), L.dptr(rec)), "write_at_path")


def read_at_path(directory, j12, m12, l, parity=0):
    """The records [n][8] of directory/AT_<j12>_<m12>_<l>_<parity>."""
    lib = L.load()
    n = ctypes.c_int()
    d = str(directory).encode()
    L.check(lib.mqct_hip_read_at_path(d, int(j12), int(m12), int(l), int(parity), 0, None, ctypes.byref(n)),
            "read_at_path")
    rec = np.zeros((n.value, 8))
    L.check(lib.mqct_hip_read_at_path(d, int(j12), int(m12), int(l), int(parity), n.value, L.dptr(rec),
                                      ctypes.byref(n)), "read_at_path")
    return rec


def at_first_step(dev, wl_small, b, E_cm, s_st=1, time_step=40.0, rec_cap=8000):
    """The first AT run: CC-MQCT trajectories on a small basis that record
    the classical path (SAVE_TRAJECT; propagator.f:3457-3462).  Returns the
    path records and the integer L of every trajectory."""
    from .matrix import MatrixBuilder
    mb = MatrixBuilder(dev, wl_small)
    mb.build("fp64", copy_out=False)
    prop = Propagator(dev, wl_small.basis, wl_small.R, mass_amu=wl_small.mass_amu, from_build=True)
    k = np.sqrt(prop.mass * 2.0 * np.asarray(E_cm) * CM_TO_HARTREE)
    l = np.floor(np.asarray(b) * k)
    rows = []
    for i in range(len(b)):
        rows += prop.trajectories([l[i]], float(E_cm[i]), s_st, time_step=time_step)
    q, st, paths = prop.run_record(rows, rec_cap)
    return paths, l, st
