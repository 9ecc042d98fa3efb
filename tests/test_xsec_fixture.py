"""Cross-section assembly against the reference's own CO+He run
(EXAMPLES_IO/SYS_TYPE_1/CO_He_dynamics, committed as data in
tests/golden/fixtures.json): from the opacities P_n(L) it printed, the
host code must reproduce its L grid, its sigma(U) (INELAST_CALC,
propagator.f:4778-4833, in angstrom^2), its billing-corrected energies and
sigma(E_coll) (propagator.f:276-287, 1208), and its J-resolved partial
cross sections.  The opacities carry 8 significant digits, so the sums
agree to ~1e-8 relative; the partial table is printed with 5."""
import json
import os

import numpy as np

from mqct_amd.basis import A_BOHR, AMUTOAU
from mqct_amd.propagate import billing_scale, inelastic_sigma, l_grid

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))["co_he"]


def _setup():
    f = FIX
    mass = f["MASS_RED"] * AMUTOAU
    op = np.array(f["opacity"])
    return f, mass, op[:, 0], op[:, 1].astype(int), op[:, 2:]


def test_l_grid_and_impact_parameters():
    f, mass, b, L, P = _setup()
    ls, (J0, J1), k = l_grid(mass, f["U"], f["B_IMPCT"], 0, 0)
    assert np.array_equal(ls, L)            # 37 trajectories, L = 0..36
    assert (J0, J1) == (0, 36)
    assert np.max(np.abs(b - L / k)) <= 5e-8 * b.max()


def test_sigma_U_and_billing_match_reference_output():
    f, mass, b, L, P = _setup()
    ls, Jr, k = l_grid(mass, f["U"], f["B_IMPCT"], 0, 0)
    sig = inelastic_sigma(P, ls, Jr, 0, k) * A_BOHR ** 2
    E_ch = f["BE"] * np.arange(f["JMAX"] + 1) * (np.arange(f["JMAX"] + 1) + 1.0)
    scale = billing_scale(E_ch, 1, f["U"])
    for ilv, flv, sU, Ecoll, sE in f["sigma"]:
        n = int(flv) - 1
        if n == 0:
            continue  # the elastic channel is printed with the elastic correction, not this sum
        assert abs(sig[n] - sU) <= 2e-7 * sU, (flv, sig[n], sU)
        E_bill = f["U"] * A_BOHR ** 2 / scale[n]
        assert abs(E_bill - Ecoll) <= 1e-6 * Ecoll
        assert abs(sig[n] * scale[n] / A_BOHR ** 2 - sE) <= 2e-7 * sE


def test_partial_cross_sections_per_J():
    f, mass, b, L, P = _setup()
    ls, Jr, k = l_grid(mass, f["U"], f["B_IMPCT"], 0, 0)
    part = np.array(f["partial_J"])
    for row in part:
        J = int(row[0])
        ref = row[1:]
        got = (2 * J + 1) * np.pi / k ** 2 * P[J] * A_BOHR ** 2  # j = 0: l = J only
        m = ref > 0
        assert np.max(np.abs(got[m] - ref[m]) / ref[m]) <= 1e-4


def test_inelastic_weights_reproduce_the_sums():
    """The per-trajectory weights the GPU kernel takes give the same sums
    as INELAST_CALC's double loop over J and l (propagator.f:4778-4833),
    for j_ini = 0 and 2 and the J-resolved rows."""
    f, mass, b, L, P = _setup()
    from mqct_amd.propagate import inelastic_weights
    for j_ini in (0, 2):
        ls, Jr, k = l_grid(mass, f["U"], f["B_IMPCT"], j_ini, j_ini)
        Pj = P[:len(ls)] if len(ls) <= len(P) else np.resize(P, (len(ls), P.shape[1]))
        Wt = inelastic_weights(ls, Jr, j_ini, k)
        assert np.allclose(Wt[0] @ Pj, inelastic_sigma(Pj, ls, Jr, j_ini, k), rtol=1e-14, atol=0)
        WJ = inelastic_weights(ls, Jr, j_ini, k, per_J=True)
        assert np.allclose(WJ.sum(0), Wt[0], rtol=1e-14, atol=0)
