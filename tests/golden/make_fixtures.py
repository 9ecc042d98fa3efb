"""Extracts the small data fixtures the CPU tests pin against from the
reference's own example outputs (run once, in the build container where the
reference tree is mounted; the GPU runs never read the reference).

Fixtures written to tests/golden/fixtures.json:
  h2o_he : EXAMPLES_IO/SYS_TYPE_4/H2O_He_external_rot_wfn -- rotational
           constants and channel list of H2O_He.inp, level energies and
           eigenvectors of USER_DEFINED_BASIS.DAT, the state table, MIJ
           index list and R grid of MTRX.DAT.
  mf_he  : EXAMPLES_IO/SYS_TYPE_4/MF_He_CS_MQCT -- constants and channel
           list of MF_He_CS_MQCT.inp, channel energies printed in
           USER_INPUT_CHECK.out, state table and MIJ index list of MTRX.DAT.
  co_he  : EXAMPLES_IO/SYS_TYPE_1/CO_He_dynamics -- the $BASIS/$SYSTEM
           constants of CO_He_dynamics.inp, the opacity table P_n(L) of
           OPACITY_FUNCTIONS_L.out (input of the cross-section assembly),
           and its outputs: sigma(U), E_coll and sigma(E_coll) of
           CROSS_SECTIONS.out and the J-resolved table of
           PARTIAL_CROSS_SECTIONS_J.out.
"""
import json
import os
import re

REF = "/root/reference/EXAMPLES_IO/SYS_TYPE_4"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")


def mtrx_tables(path):
    lines = open(path).read().splitlines()
    hdr = {}
    for ln in lines[:6]:
        m = re.match(r"\s*([A-Z_]+)=\s*(\d+)", ln)
        if m:
            hdr[m.group(1)] = int(m.group(2))
    i = next(k for k, ln in enumerate(lines) if "#STATE" in ln)
    ns = hdr["NUMBER_OF_STATES"]
    states = [[int(x) for x in lines[i + 1 + s].split()] for s in range(ns)]
    j = next(k for k, ln in enumerate(lines) if "MIJ_INDEX" in ln)
    T = hdr["MATRIX_SIZE"]
    pairs = [[int(x) for x in lines[j + 1 + k].split()] for k in range(T)]
    r = next(k for k, ln in enumerate(lines) if "R_COM" in ln)
    nr = hdr["R_GRID_SIZE"]
    R = [float(lines[r + 1 + k].split()[1]) for k in range(nr)]
    return hdr, states, pairs, R


def h2o_he():
    d = os.path.join(REF, "H2O_He_external_rot_wfn")
    inp = open(os.path.join(d, "H2O_He.inp")).read()
    A = float(re.search(r"\bA = ([\d.]+)", inp).group(1))
    B = float(re.search(r"\bB = ([\d.]+)", inp).group(1))
    C = float(re.search(r"\bC = ([\d.]+)", inp).group(1))
    basis = open(os.path.join(d, "USER_DEFINED_BASIS.DAT")).read().splitlines()
    chans, vecs = [], []
    k = 1
    while not basis[k].startswith("Expansion"):
        n, j, ka, kc, e = basis[k].split()
        chans.append([int(j), int(ka), int(kc), float(e)])
        k += 1
    k += 1
    for c in range(len(chans)):
        parts = basis[k + c].split()
        vecs.append([float(x) for x in parts[1:]])
    hdr, states, pairs, R = mtrx_tables(os.path.join(d, "MTRX.DAT"))
    return {"A": A, "B": B, "C": C, "channels": chans, "eigvecs": vecs, "header": hdr,
            "states": states, "pairs": pairs, "R": R}


def mf_he():
    d = os.path.join(REF, "MF_He_CS_MQCT")
    inp = open(os.path.join(d, "MF_He_CS_MQCT.inp")).read()
    A = float(re.search(r"\bA = ([\d.]+)", inp).group(1))
    B = float(re.search(r"\bB = ([\d.]+)", inp).group(1))
    C = float(re.search(r"\bC = ([\d.]+)", inp).group(1))
    chk = open(os.path.join(d, "USER_INPUT_CHECK.out")).read().splitlines()
    i = next(k for k, ln in enumerate(chk) if "CHANNELS ENERGIES" in ln)
    chans = []
    for ln in chk[i + 2:]:
        p = ln.split()
        if len(p) != 5 or not p[0].isdigit():
            break
        chans.append([int(p[1]), int(p[2]), int(p[3]), float(p[4])])
    hdr, states, pairs, R = mtrx_tables(os.path.join(d, "MTRX.DAT"))
    return {"A": A, "B": B, "C": C, "channels": chans, "header": hdr, "states": states,
            "pairs": pairs, "n_r": len(R)}


def _table(lines, start):
    rows = []
    for ln in lines[start:]:
        p = ln.split()
        if not p or not re.match(r"^-?\d", p[0]):
            break
        rows.append([float(x) for x in p])
    return rows


def co_he():
    d = "/root/reference/EXAMPLES_IO/SYS_TYPE_1/CO_He_dynamics"
    inp = open(os.path.join(d, "CO_He_dynamics.inp")).read()
    num = lambda key: float(re.search(r"\b%s = ([\d.]+)" % key, inp).group(1))
    op = open(os.path.join(d, "OPACITY_FUNCTIONS_L.out")).read().splitlines()
    i = next(k for k, ln in enumerate(op) if ln.startswith("B_IMPACT"))
    opac = _table(op, i + 1)
    cs = open(os.path.join(d, "CROSS_SECTIONS.out")).read().splitlines()
    i = next(k for k, ln in enumerate(cs) if "sigma(U)" in ln)
    sig = _table(cs, i + 1)
    pj = open(os.path.join(d, "PARTIAL_CROSS_SECTIONS_J.out")).read().splitlines()
    i = next(k for k, ln in enumerate(pj) if ln.startswith("J_TOTAL"))
    part = _table(pj, i + 1)
    return {"BE": num("BE"), "JMAX": int(num("JMAX")), "MASS_RED": num("MASS_RED"), "U": num("U_ENERGY"),
            "B_IMPCT": num("B_IMPCT"), "RMAX": num("RMAX"),
            "opacity": opac,            # [b, L, P_1..P_21]
            "sigma": sig,               # [ilv, flv, sigma(U) A^2, E_coll cm^-1, sigma(E_coll) A^2]
            "partial_J": part}          # [J, sigma_J per channel]


if __name__ == "__main__":
    json.dump({"h2o_he": h2o_he(), "mf_he": mf_he(), "co_he": co_he()}, open(OUT, "w"), indent=1)
    print("wrote", OUT)
