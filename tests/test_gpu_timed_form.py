"""The benchmark's timed form against the oracle: AT-MQCT with odeint
(Cash-Karp, EPS_ODEINT 1e-8, MIN_TMSTP 20 a.u.) at TIME_STEP 10 (the timed
form: first trials at or below MIN_TMSTP, steps accepted at the floor) and
at TIME_STEP 60 (every first trial above MIN_TMSTP: rkqs's error control
decides each step) on the full asymmetric top + atom system (1771 states, 86 779 pairs, 300 knots,
MIJ_SHIFT on), for 8 trajectories of the 4096-trajectory batch with paths
from the first AT step.

* batched=1 (FP64 knots and products) against the oracle's DOTRAJECT:
  amplitudes within 1e-6 (odeint's step decisions may round differently);
* batched=3 (FP32 knots, FP32 products in the pipelined GEMM forms: the
  timed form): amplitudes within 1e-6 and the sample's cross sections
  (INELAST weights of each trajectory's l and E) within 1e-5 relative.

The oracle runs on 8 host processes (about 20 s of CPU per trajectory)."""
import os
import tempfile
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.basis import AMUTOAU, CM_TO_HARTREE

pytestmark = pytest.mark.gpu

PICK = [3, 517, 1029, 1555, 2048, 2600, 3333, 4090]
HMIN, EPS = 20.0, 1e-8


def _oracle_traj(args):
    """One trajectory of the oracle in a worker process (spawned: no GPU)."""
    npz, path, row, mass = args
    import oracle.oracle as O
    z = np.load(npz)
    dyn = O.DynOracle(z["j12"], z["m12"], z["Ej"], z["ind"], z["R"], z["M"], z["D"], mass, path=path)
    q, st, _ = dyn.dotraject(L.MQCT_PROP_AT_ODEINT, 1, row["l_real"], row["E_coll"], row["dt"], eps=row["eps"],
                             hmin=row["hmin"], R_start=row["R_start"])
    return q, st


@pytest.mark.parametrize("time_step", [10.0, 60.0])
def test_timed_form_matches_oracle(dev, tmp_path, time_step):
    import oracle.oracle as O
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import Propagator, at_first_step
    bb, E = W.at_batch(4096)
    bb, E = bb[PICK], E[PICK]
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=2), bb, E, time_step=40.0, rec_cap=6000)
    wl = W.asym_top_atom()
    b = wl.basis
    M, D, S = MatrixBuilder(dev, wl).build("fp64")
    prop = Propagator(dev, b, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(PICK)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=time_step, eps=EPS, hmin=HMIN, at_paths=[i])
    q1, s1 = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=1)
    q3, s3 = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=3)
    rej = (s1["n_rhs"] - 6 * s1["n_steps"]) / 5
    print("steps", s1["n_steps"].tolist(), "rejected trials", rej.astype(int).tolist(), flush=True)
    if time_step > 20.0:
        assert np.all(rej > 0)  # rkqs's error control runs
    else:
        assert np.all(rej == 0)  # every step at the floor
    npz = os.path.join(str(tmp_path), "m.npz")
    np.savez(npz, j12=b.j12, m12=b.m12, Ej=b.Ej * CM_TO_HARTREE, ind=b.ind_mat, R=wl.R, M=O.mij_shift(M), D=D)
    ctx = mp.get_context("spawn")
    with ProcessPoolExecutor(max_workers=8, mp_context=ctx) as ex:
        res = list(ex.map(_oracle_traj, [(npz, paths[i], rows[i], wl.mass_amu * AMUTOAU) for i in range(len(PICK))]))
    N = b.n_states
    qo = np.array([r[0][:2 * N] for r in res])
    d1 = np.max(np.abs(q1[:, :2 * N] - qo))
    d3 = np.max(np.abs(q3[:, :2 * N] - qo))
    k2 = prop.mass * 2.0 * E * CM_TO_HARTREE
    w = (2.0 * l + 1.0) * np.pi / k2
    so = w @ prop.opacities(np.hstack([qo, np.zeros((len(PICK), 8))]))
    s3x = w @ prop.opacities(q3)
    m = so >= 1e-6 * so.max()
    xs_err = np.max(np.abs(s3x[m] - so[m]) / so[m])
    print(f"batched=1 amp diff {d1:.2e}; batched=3 amp diff {d3:.2e}, xsec err {xs_err:.2e} over {m.sum()} channels")
    assert d1 <= 1e-6
    assert d3 <= 1e-6
    assert xs_err <= 1e-5
