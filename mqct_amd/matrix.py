"""Matrix build through the HIP library: the host-side mirror of MATRIX_MAKE's
compute branch (matrix_ini.f:1916-2016) for a whole basis in one call.

    with Device() as dev:
        mb = MatrixBuilder(dev, workload)
        mat_el, mat_el_der, skipped = mb.build(precision="fp64")

Arrays follow the reference: mat_el[k, i] is Mat_el(i, k) (element k, R
point i), Hartree.
"""
import ctypes

import numpy as np

from . import _lib as L
from .basis import AngularGrid, Basis

PRECISIONS = {"fp64": L.MQCT_PREC_FP64, "fp32": L.MQCT_PREC_FP32,
              "bf16x3": L.MQCT_PREC_BF16X3}


class Device:
    """One library handle on one GPU."""

    def __init__(self, device=0):
        self.lib = L.load()
        h = ctypes.c_void_p()
        L.check(self.lib.mqct_hip_create(device, ctypes.byref(h)), "mqct_hip_create")
        self.h = h

    def set_mij_shift(self, on=True):
        """MIJ_SHIFT= (default YES): propagate M(R) - M(R_max)."""
        L.check(self.lib.mqct_hip_set_mij_shift(self.h, int(bool(on))), "set_mij_shift")

    def close(self):
        if self.h:
            self.lib.mqct_hip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grid_struct(g: AngularGrid):
    keep = []

    def arr(a, n):
        if a is None:
            a = np.zeros(n) if n == 1 else None
        a = L.as_f64(a if a is not None else np.zeros(n))
        keep.append(a)
        return L.dptr(a)

    one = np.ones(1)
    s = L.GridT()
    s.n_beta, s.xb, s.wb, s.xbm, s.xbl = g.n_beta, arr(g.xb, g.n_beta), arr(g.wb, g.n_beta), g.xbm, g.xbl
    s.n_gamma, s.xg, s.xgm, s.xgl = g.n_gamma, arr(g.xg, g.n_gamma), g.xgm, g.xgl
    s.wg = arr(g.wg if g.wg is not None else one, g.n_gamma)
    s.n_alpha2, s.xaa, s.xaam, s.xaal = g.n_alpha2, arr(g.xaa, g.n_alpha2), g.xaam, g.xaal
    s.waa = arr(g.waa if g.waa is not None else one, g.n_alpha2)
    s.n_beta2, s.xbb, s.xbbm, s.xbbl = g.n_beta2, arr(g.xbb, g.n_beta2), g.xbbm, g.xbbl
    s.wbb = arr(g.wbb if g.wbb is not None else one, g.n_beta2)
    s.n_gamma2, s.xgg, s.xggm, s.xggl = g.n_gamma2, arr(g.xgg, g.n_gamma2), g.xggm, g.xggl
    s.wgg = arr(g.wgg if g.wgg is not None else one, g.n_gamma2)
    s.eq_gamma, s.ref_xz = g.eq_gamma, g.ref_xz
    return s, keep


class MatrixBuilder:
    def __init__(self, dev: Device, wl, vgrid=None):
        """wl: a workloads.Workload (basis, grid, R, potential).  vgrid, if
        given, is the host V grid in POT_STORE order (E_UNITS); otherwise the
        library evaluates the workload's synthetic potential on the GPU."""
        self.dev, self.lib, self.h = dev, dev.lib, dev.h
        self.wl = wl
        b: Basis = wl.basis
        self._keep = []
        j12, m12, ich = L.as_i32(b.j12), L.as_i32(b.m12), L.as_i32(b.indx_chann)
        j1, j2 = L.as_i32(b.j1_ch), (L.as_i32(b.j2_ch) if b.j2_ch is not None else None)
        e1 = L.as_f64(b.eig1) if b.eig1 is not None else None
        e2 = L.as_f64(b.eig2) if b.eig2 is not None else None
        L.check(self.lib.mqct_hip_set_basis(self.h, b.coll_type, b.n_states, L.iptr(j12), L.iptr(m12),
                                            L.iptr(ich), b.n_channels, L.iptr(j1), L.iptr(j2),
                                            b.ldvec, L.dptr(e1), L.dptr(e2)), "set_basis")
        if getattr(b, "identical", False):
            same = ((b.j1_ch == b.j2_ch) & (b.ka1_ch == b.ka2_ch) & (b.kc1_ch == b.kc2_ch)).astype(np.int32)
            ps, pb = L.as_i32(b.parity_state), L.as_i32(b.parity_state_bk)
            self._keep += [ps, pb, same]
            L.check(self.lib.mqct_hip_set_identical(self.h, 1, L.iptr(ps), L.iptr(pb), L.iptr(same)),
                    "set_identical")
        gs, keep = grid_struct(wl.grid)
        L.check(self.lib.mqct_hip_set_grid(self.h, ctypes.byref(gs)), "set_grid")
        R = L.as_f64(wl.R)
        self.R = R
        if vgrid is not None:
            V = L.as_f64(vgrid)
            L.check(self.lib.mqct_hip_set_vgrid(self.h, len(R), L.dptr(R), L.dptr(V)), "set_vgrid")
        else:
            par = L.as_f64(wl.pes_par)
            L.check(self.lib.mqct_hip_eval_pes(self.h, len(R), L.dptr(R), wl.pes_kind, len(par),
                                               L.dptr(par)), "eval_pes")
        self.ind_mat = L.as_i32(b.ind_mat)

    def build(self, precision="fp64", copy_out=True):
        prec = PRECISIONS[precision] if isinstance(precision, str) else precision
        b = self.wl.basis
        T, n = b.total_size, len(self.R)
        ir1, ir2 = self.wl.ir_cut
        if copy_out:
            M = np.zeros((T, n))
            D = np.zeros((T, n))
            S = np.zeros(T, np.uint8)
            L.check(self.lib.mqct_hip_matrix_build(self.h, prec, T, L.iptr(self.ind_mat), self.wl.conv_unit_e,
                                                   self.wl.mij_zero_cut, ir1, ir2, L.dptr(M), L.dptr(D),
                                                   S.ctypes.data_as(L.c_u8_p)), "matrix_build")
            return M, D, S
        L.check(self.lib.mqct_hip_matrix_build(self.h, prec, T, L.iptr(self.ind_mat), self.wl.conv_unit_e,
                                               self.wl.mij_zero_cut, ir1, ir2, None, None, None),
                "matrix_build")
        return None

    def timing(self):
        t = np.zeros(5)
        L.check(self.lib.mqct_hip_build_timing(self.h, L.dptr(t)), "build_timing")
        return {"psi_ms": t[0], "weights_ms": t[1], "contract_ms": t[2], "post_ms": t[3], "total_ms": t[4]}


def write_mtrx_uf(path, wl, mat_el):
    lib = L.load()
    b = wl.basis
    qn = b.chan_qn()
    L.check(lib.mqct_hip_write_mtrx_uf(path.encode(), b.coll_type, b.n_channels, b.n_states, b.total_size,
                                       len(wl.R), L.iptr(L.as_i32(b.indx_chann)), L.iptr(L.as_i32(b.j12)),
                                       L.iptr(L.as_i32(b.m12)), L.iptr(qn), L.iptr(L.as_i32(b.ind_mat)),
                                       L.dptr(L.as_f64(wl.R)), L.dptr(L.as_f64(mat_el))), "write_mtrx_uf")


def read_mtrx_uf(path):
    lib = L.load()
    hdr = np.zeros(5, np.int64)
    L.check(lib.mqct_hip_read_mtrx_uf(path.encode(), hdr.ctypes.data_as(L.c_i64_p), None, None, None, None,
                                      None, None, None), "read_mtrx_uf header")
    coll, nch, ns, T, nr = (int(x) for x in hdr)
    nq = {1: 1, 4: 3, 0: 6}[coll]
    ich, j12, m12 = (np.zeros(ns, np.int32) for _ in range(3))
    qn = np.zeros((nch, nq), np.int32)
    ind = np.zeros((T, 2), np.int32)
    R = np.zeros(nr)
    M = np.zeros((T, nr))
    L.check(lib.mqct_hip_read_mtrx_uf(path.encode(), hdr.ctypes.data_as(L.c_i64_p), L.iptr(ich), L.iptr(j12),
                                      L.iptr(m12), L.iptr(qn), L.iptr(ind), L.dptr(R), L.dptr(M)),
            "read_mtrx_uf")
    return {"coll_type": coll, "n_channels": nch, "n_states": ns, "total_size": T, "n_r": nr,
            "indx_chann": ich, "j12": j12, "m12": m12, "chan_qn": qn, "ind_mat": ind, "R": R, "mat_el": M}


def write_mij_dir(path, wl, mat_el, n_files, unformatted=False, parity_state=None):
    """The parallel-IO matrix directory (MATRIX_TRUNCATED) of a split build:
    MATRIX_INFO.DAT, then the elements above MIJ_CUTOFF in n_files
    MIJ_/Ind_ file pairs with MATRIX_NONZERO_INDEX.DAT and
    MATRIX_FILE_INDEX.DAT (matrix_ini.f:1545-1575, 6836-7244, 9053-9145)."""
    import os
    lib = L.load()
    b = wl.basis
    os.makedirs(path, exist_ok=True)
    ps = L.as_i32(parity_state) if parity_state is not None else (
        L.as_i32(b.parity_state) if getattr(b, "identical", False) else None)
    L.check(lib.mqct_hip_write_mij_info(path.encode(), int(unformatted), b.coll_type, b.n_channels, b.n_states,
                                        b.total_size, len(wl.R), L.iptr(L.as_i32(b.indx_chann)),
                                        L.iptr(L.as_i32(b.j12)), L.iptr(L.as_i32(b.m12)), L.iptr(b.chan_qn()),
                                        L.iptr(ps), L.iptr(L.as_i32(b.ind_mat)), L.dptr(L.as_f64(wl.R))),
            "write_mij_info")
    ir1, ir2 = wl.ir_cut
    M = L.as_f64(mat_el)
    L.check(lib.mqct_hip_write_mij_dir(path.encode(), int(unformatted), n_files, b.total_size, len(wl.R),
                                       L.iptr(L.as_i32(b.ind_mat)), L.dptr(M), wl.mij_zero_cut, ir1, ir2),
            "write_mij_dir")


def read_mij_dir(path, n_r, unformatted=False):
    """bk_read_matrix for one reader: (ind_mat[n_nonzero][2], mat_el
    [n_nonzero][n_r], the file's cut-off in cm^-1), the elements in file
    order; mij_spline then applies MIJ_SHIFT and the spline."""
    lib = L.load()
    nnz = np.zeros(1, np.int64)
    cut = np.zeros(1)
    L.check(lib.mqct_hip_read_mij_dir(path.encode(), int(unformatted), n_r, 0, nnz.ctypes.data_as(L.c_i64_p),
                                      L.dptr(cut), None, None), "read_mij_dir sizes")
    n = int(nnz[0])
    ind = np.zeros((n, 2), np.int32)
    M = np.zeros((n, n_r))
    L.check(lib.mqct_hip_read_mij_dir(path.encode(), int(unformatted), n_r, n, nnz.ctypes.data_as(L.c_i64_p),
                                      L.dptr(cut), L.iptr(ind), L.dptr(M)), "read_mij_dir")
    return ind, M, float(cut[0])


def mij_spline(dev, R, mat_el):
    """READ_MIJ_USER's post-read step (matrix_ini.f:6603-6621) on the GPU:
    MIJ_SHIFT per the handle's setting, then the clamped spline.  Returns
    (Mat_el, Mat_el_der) ready for Propagator(mat_el=..., mat_el_der=...)."""
    M = np.array(mat_el, dtype=np.float64, order="C", copy=True)
    D = np.zeros_like(M)
    R = L.as_f64(R)
    L.check(dev.lib.mqct_hip_mij_spline(dev.h, len(R), L.dptr(R), M.shape[0], L.dptr(M), L.dptr(D)),
            "mij_spline")
    return M, D


def read_mij_user(dev, path):
    """MTRX_UF.DAT -> the arrays READ_MIJ_USER leaves for the propagation
    (matrix_ini.f:5794-6621): the file's tables plus the shifted and
    re-splined Mat_el / Mat_el_der."""
    f = read_mtrx_uf(path)
    f["mat_el"], f["mat_el_der"] = mij_spline(dev, f["R"], f["mat_el"])
    return f


def _main(argv=None):
    """Times the matrix build of a named workload on GPU 0 in each precision
    form (twice each: the first build of a form also times its contraction
    variants, which the reported times leave out):

        python -m mqct_amd.matrix two_asym_tops fp64,fp32,bf16x3
    """
    import argparse
    from . import workloads as W
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("workload", choices=("diatom_atom", "asym_top_atom", "two_asym_tops", "two_asym_tops_c3"))
    ap.add_argument("precisions", nargs="?", default="fp64,fp32,bf16x3")
    ap.add_argument("--reps", type=int, default=2)
    a = ap.parse_args(argv)
    wl = getattr(W, a.workload)()
    with Device(0) as dev:
        mb = MatrixBuilder(dev, wl)
        for prec in a.precisions.split(","):
            for rep in range(a.reps):
                mb.build(prec, copy_out=False)
                v, kc, nw, nc = (ctypes.c_int() for _ in range(4))
                L.check(dev.lib.mqct_hip_contract_variant(dev.h, PRECISIONS[prec], ctypes.byref(v), ctypes.byref(kc),
                                                          ctypes.byref(nw), ctypes.byref(nc)), "contract_variant")
                t = mb.timing()
                print("%s %s build %d: contraction variant %d (%d grid points per stage, %d waves); %s" % (
                    a.workload, prec, rep, v.value, kc.value, nw.value,
                    ", ".join("%s %.2f" % (k, float(x)) for k, x in t.items())), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(_main())
