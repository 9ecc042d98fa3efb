"""Writes tests/golden/h2ohe_vgrid.json: the reference's own H2O-He potential
(EXAMPLES_IO/SYS_TYPE_4/H2O_He_external_rot_wfn/PES_H2OHe, compiled from its
source by `make -C oracle ref` into oracle/_ref/) evaluated on that example's
SYS_TYPE 4 grid, and the matrix elements the example's MTRX.DAT holds.

Run once, in the build container where the reference tree is mounted; the
tests read only the JSON.

The deck (H2O_He.inp): GRD_ANG2 = 10 Gauss-Legendre beta points on [0, pi],
GRD_ANG3 = 20 equidistant gamma points (the default equidistant gamma grid,
pot_comp.f:596-610), GRD_R = 2 points R = 6 and 8 bohr (RMIN, RMAX;
R_UNITS = A.U.), E_UNITS = CM-1.  The grid is filled as INI_POT_STORAGE does
(pot_comp.f:613-631): V_3_1(i3, i2, i) = PES(R, -beta, -gamma) through
V_POT_TOP_ATOM (pes_sys_type.f:32-37).

Output: R, beta, gamma, V[R][beta][gamma] (cm^-1), and mij[k][iR] (Hartree,
the MTRX.DAT block "#k_matel #iR Mij(iR,k_matel)", 12 significant digits).
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EX = "/root/reference/EXAMPLES_IO/SYS_TYPE_4/H2O_He_external_rot_wfn"
EXE = os.path.join(ROOT, "oracle", "_ref", "h2ohe", "ref_pes_h2ohe")
OUT = os.path.join(HERE, "h2ohe_vgrid.json")


def mtrx_mij(path):
    lines = open(path).read().splitlines()
    i = next(k for k, ln in enumerate(lines) if "Mij(iR,k_matel)" in ln)
    rows = [ln.split() for ln in lines[i + 1:] if ln.strip()]
    T = max(int(r[0]) for r in rows)
    nr = max(int(r[1]) for r in rows)
    M = np.zeros((T, nr))
    for k, ir, v in rows:
        M[int(k) - 1, int(ir) - 1] = float(v)
    return M


def main():
    import sys
    sys.path.insert(0, ROOT)
    from mqct_amd import basis as B
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    g = B.grid_type4(10, 20)
    beta = g.xbm + g.xbl * np.asarray(g.xb)
    gamma = np.asarray(g.xg)
    R = np.array([6.0, 8.0])
    pts = [(r, -b, -c) for r in R for b in beta for c in gamma]
    inp = "%d\n" % len(pts) + "".join("%.17g %.17g %.17g\n" % p for p in pts)
    with tempfile.TemporaryDirectory() as wd:
        # the potential opens its data files as PES_H2OHe/<name>
        os.symlink(os.path.join(EX, "PES_H2OHe"), os.path.join(wd, "PES_H2OHe"))
        r = subprocess.run([EXE], input=inp, capture_output=True, text=True, cwd=wd, check=True)
    V = np.array([float(x) for x in r.stdout.split()]).reshape(len(R), len(beta), len(gamma))
    mij = mtrx_mij(os.path.join(EX, "MTRX.DAT"))
    json.dump({"R": R.tolist(), "beta": beta.tolist(), "gamma": gamma.tolist(), "V": V.tolist(),
               "mij": mij.tolist(), "e_units": "cm-1", "mij_units": "hartree",
               "grid": {"n_beta": 10, "n_gamma": 20}}, open(OUT, "w"))
    print("wrote", OUT, "V range %.3f .. %.3f cm^-1" % (V.min(), V.max()))


if __name__ == "__main__":
    main()
