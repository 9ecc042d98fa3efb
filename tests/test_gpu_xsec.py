"""Cross sections on the GPU (mqct_hip_cross_sections: DOTRAJECT's channel
probabilities and INELAST_CALC's sums, propagator.f:4364-4383, 4752-4983).

* Against the reference's own CO+He run (EXAMPLES_IO/SYS_TYPE_1/
  CO_He_dynamics, data in tests/golden/fixtures.json): amplitudes that
  carry its printed opacities P_n(L) give its sigma(U) (2e-7, the
  opacities' 8 digits) and its J-resolved table (1e-4, printed with 5).
* Against a float64 host evaluation of the same sums on random amplitudes
  of an uneven basis (1e-13 relative), with the opacities bit-identical to
  a fixed-order host sum, two calls bit-identical to each other, and a
  malformed channel table refused.
"""
import json
import os

import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import basis as B
from mqct_amd import workloads as W
from mqct_amd.basis import A_BOHR, AMUTOAU
from mqct_amd.propagate import billing_scale, gpu_cross_sections, inelastic_weights, l_grid

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))["co_he"]


def test_co_he_reference_cross_sections(dev):
    f = FIX
    mass = f["MASS_RED"] * AMUTOAU
    op = np.array(f["opacity"])
    L_, P = op[:, 1].astype(int), op[:, 2:]
    ls, Jr, k = l_grid(mass, f["U"], f["B_IMPCT"], 0, 0)
    assert np.array_equal(ls, L_)
    b = B.ini_arrays(B.channels_type1(0, f["JMAX"], f["BE"]))
    N = b.n_states
    # the probability of channel n in its m = 0 state
    q = np.zeros((len(ls), 2 * N + 8))
    for n in range(b.n_channels):
        q[:, b.chann_indx[n] - 1] = np.sqrt(P[:, n])
    Wt = inelastic_weights(ls, Jr, 0, k)
    sig, Pg = gpu_cross_sections(dev, q, b, Wt, opacity=True)
    assert np.max(np.abs(Pg - P)) <= 1e-15
    sig = sig[0] * A_BOHR ** 2
    scale = billing_scale(b.E_ch, 1, f["U"])
    for ilv, flv, sU, Ecoll, sE in f["sigma"]:
        n = int(flv) - 1
        if n == 0:
            continue  # printed with the elastic correction, not this sum
        assert abs(sig[n] - sU) <= 2e-7 * sU, (flv, sig[n], sU)
        assert abs(sig[n] * scale[n] / A_BOHR ** 2 - sE) <= 2e-7 * sE
    WJ = inelastic_weights(ls, Jr, 0, k, per_J=True)
    sJ = gpu_cross_sections(dev, q, b, WJ) * A_BOHR ** 2
    for row in np.array(f["partial_J"]):
        J = int(row[0])
        m = row[1:] > 0
        assert np.max(np.abs(sJ[J - Jr[0]][m] - row[1:][m]) / row[1:][m]) <= 1e-4


def test_random_amplitudes_against_float64_sums(dev):
    b = W.asym_top_atom(jmax=6, n_r=4).basis        # channels of 1..13 states
    rng = np.random.default_rng(5)
    T, N = 1000, b.n_states
    q = rng.standard_normal((T, 2 * N + 8))
    Wt = rng.uniform(0.0, 1.0, (3, T))
    sig, P = gpu_cross_sections(dev, q, b, Wt, opacity=True)
    a2 = q[:, :N] ** 2 + q[:, N:2 * N] ** 2
    Pref = np.zeros((T, b.n_channels))
    for s in range(N):                                # states ascending, as the kernel
        Pref[:, b.indx_chann[s] - 1] += a2[:, s]
    assert np.max(np.abs(P - Pref) / Pref) <= 1e-15
    ref = Wt @ Pref
    assert np.max(np.abs(sig - ref) / ref) <= 1e-13
    sig2 = gpu_cross_sections(dev, q, b, Wt)
    assert np.array_equal(sig, sig2)


def test_malformed_channel_table_is_refused(dev):
    b = W.asym_top_atom(jmax=2, n_r=4).basis
    q = np.ones((4, 2 * b.n_states + 8))
    bad = b.indx_chann.copy()
    bad[[0, -1]] = bad[[-1, 0]]                       # not channel-major
    with pytest.raises(L.MqctError, match="INI_ARRAYS order"):
        L.check(dev.lib.mqct_hip_cross_sections(dev.h, 4, b.n_states, q.shape[1], L.dptr(q), b.n_channels,
                                                L.iptr(L.as_i32(bad)), 1, L.dptr(np.ones(4)),
                                                L.dptr(np.zeros(b.n_channels)), None), "cross_sections")
