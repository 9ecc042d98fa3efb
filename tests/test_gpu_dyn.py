"""GPU parity of the propagation (the HIP library through its C ABI)
against the CPU oracle's DOTRAJECT restatement.

The matrix (Mat_el, Mat_el_der) is the oracle's FP64 build of the
diatom+atom workload, so both sides integrate the same equations.  Bars:
fixed-step RK4 (same step sequence on both sides): amplitudes within 1e-9
absolute, classical variables within 1e-9 relative.  Adaptive odeint: the
step-size decisions may round differently, so amplitudes within 1e-6
(the tolerance is 1e-8 per step).  Norm drift is reported.
"""
import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.basis import CM_TO_HARTREE, AMUTOAU
from mqct_amd.propagate import Propagator
import oracle.oracle as O

pytestmark = pytest.mark.gpu

U_CM = 500.0


@pytest.fixture(scope="module")
def c0():
    wl = W.diatom_atom()
    orc = O.oracle_for_workload(wl)
    M, D, S = orc.matrix_make(wl.basis.ind_mat, wl.conv_unit_e, wl.mij_zero_cut, *wl.ir_cut)
    b = wl.basis
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, M, D, wl.mass_amu * AMUTOAU)
    return wl, M, D, dyn


def compare(q_gpu, q_ora, N, amp_tol, cl_tol):
    a = np.max(np.abs(q_gpu[:2 * N] - q_ora[:2 * N]))
    c = np.max(np.abs(q_gpu[2 * N:2 * N + 6] - q_ora[2 * N:2 * N + 6]) /
               np.maximum(np.abs(q_ora[2 * N:2 * N + 6]), 1e-6))
    assert a <= amp_tol, f"amplitude diff {a}"
    assert c <= cl_tol, f"classical diff {c}"
    return a, c


@pytest.mark.parametrize("mode", [L.MQCT_PROP_CC_RK4, L.MQCT_PROP_CC_ODEINT])
def test_cc_trajectories_match_oracle(dev, c0, mode):
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    assert len(ls) == 83  # L = 0..82 at U=500 cm^-1, mu=3.6 amu, b<=15
    pick = [0, 23, 47, 71, 82]
    rows = prop.trajectories([ls[i] for i in pick], U_CM, s_st=1, time_step=10.0, eps=1e-8, hmin=1e-4)
    q, st = prop.run(mode, rows)
    N = b.n_states
    tol = (1e-9, 1e-9) if mode == L.MQCT_PROP_CC_RK4 else (1e-6, 1e-6)
    for i, r in enumerate(rows):
        qo, so, _ = dyn.dotraject(mode, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"], hmin=r["hmin"],
                                  R_start=r["R_start"])
        assert st["fail"][i] == so["fail"] == 0
        assert st["n_steps"][i] == so["n_steps"] or mode == L.MQCT_PROP_CC_ODEINT
        compare(q[i], qo, N, *tol)
        assert st["norm_err"][i] < 1e-6


@pytest.mark.parametrize("mode", [L.MQCT_PROP_AT_RK4, L.MQCT_PROP_AT_ODEINT])
def test_at_trajectories_match_oracle(dev, c0, mode):
    """AT-MQCT on paths recorded by the oracle's CC run (the first AT step)."""
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    pick = [5, 40, 80]
    paths = []
    for i in pick:
        r = prop.trajectories([ls[i]], U_CM, s_st=1, time_step=10.0)[0]
        _, _, rec = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, r["l_real"], r["E_coll"], r["dt"],
                                  R_start=r["R_start"], rec_cap=20000)
        paths.append(rec)
    prop.load_paths(paths)
    rows = prop.trajectories([ls[i] for i in pick], U_CM, s_st=1, time_step=10.0, eps=1e-8, hmin=1e-4,
                             at_paths=list(range(len(pick))))
    q, st = prop.run(mode, rows)
    N = b.n_states
    tol = 1e-9 if mode == L.MQCT_PROP_AT_RK4 else 1e-6
    for i, r in enumerate(rows):
        dyn.set_path(paths[i])
        qo, so, _ = dyn.dotraject(mode, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"], hmin=r["hmin"],
                                  R_start=r["R_start"])
        a = np.max(np.abs(q[i, :2 * N] - qo[:2 * N]))
        assert a <= tol, f"amplitude diff {a}"


def test_recorded_paths_match_oracle(dev, c0):
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    rows = prop.trajectories([ls[30]], U_CM, s_st=1, time_step=10.0)
    q, st, recs = prop.run_record(rows, 20000)
    _, _, rec_o = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, rows[0]["l_real"], rows[0]["E_coll"], rows[0]["dt"],
                                R_start=rows[0]["R_start"], rec_cap=20000)
    assert recs[0].shape == rec_o.shape
    assert np.max(np.abs(recs[0][:, :7] - rec_o[:, :7]) / np.maximum(np.abs(rec_o[:, :7]), 1e-6)) < 1e-9
    assert np.max(np.abs(recs[0][:, 7] - rec_o[:, 7])) < 1e-9 * np.abs(rec_o[:, 7]).max()


@pytest.mark.parametrize("mode", [L.MQCT_PROP_AT_RK4, L.MQCT_PROP_AT_ODEINT])
def test_at_batched_matches_oracle(dev, c0, mode):
    """The interval-bucketed batched form (all trajectories step together,
    spline matrices applied per knot as GEMMs) against the oracle, on paths
    of uneven length.  Same bars as the plain form."""
    wl, M, D, dyn = c0
    b = wl.basis
    prop = Propagator(dev, b, wl.R, M, D, mass_amu=wl.mass_amu)
    ls, Jr, k = prop.l_grid(U_CM, 15.0, 1)
    pick = [2, 33, 61, 80]
    paths = []
    for i in pick:
        r = prop.trajectories([ls[i]], U_CM, s_st=1, time_step=10.0)[0]
        _, _, rec = dyn.dotraject(L.MQCT_PROP_CC_RK4, 1, r["l_real"], r["E_coll"], r["dt"],
                                  R_start=r["R_start"], rec_cap=20000)
        paths.append(rec)
    prop.load_paths(paths)
    rows = prop.trajectories([ls[i] for i in pick], U_CM, s_st=1, time_step=10.0, eps=1e-8, hmin=1e-4,
                             at_paths=list(range(len(pick))))
    q, st = prop.run(mode, rows, batched=True)
    N = b.n_states
    tol = 1e-9 if mode == L.MQCT_PROP_AT_RK4 else 1e-6
    for i, r in enumerate(rows):
        dyn.set_path(paths[i])
        qo, so, _ = dyn.dotraject(mode, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"], hmin=r["hmin"],
                                  R_start=r["R_start"])
        a = np.max(np.abs(q[i, :2 * N] - qo[:2 * N]))
        assert a <= tol, f"amplitude diff {a}"
        if mode == L.MQCT_PROP_AT_RK4:
            assert st["n_steps"][i] == so["n_steps"]
        assert abs(st["t_end"][i] - so["t_end"]) <= 1e-9 * so["t_end"]
        assert st["norm_err"][i] < 1e-6


@pytest.mark.parametrize("mode", [L.MQCT_PROP_AT_RK4, L.MQCT_PROP_AT_ODEINT])
def test_at_batched_matches_plain_asym_top(dev, mode):
    """Asymmetric top + atom (j <= 3: 16 channels, 84 states in 7 m-blocks),
    48 trajectories of uneven length, several per spline interval: the
    batched form against the plain per-trajectory kernel (itself checked
    against the oracle above).  MIN_TMSTP = 20 a.u.: odeint's relative
    error control otherwise chases amplitudes at rounding-noise level
    (|y| ~ 1e-33 against yscal's 1e-30 floor) down to the minimum step and
    exhausts MAXSTP = 10000 (then both forms end with one long rk4 step and
    their results mean nothing)."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=120)
    bb, E = W.at_batch(48, seed=7)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    assert max(len(p) for p in paths) < 20000
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, eps=1e-8, hmin=20.0, at_paths=[i])
    q0, s0 = prop.run(mode, rows, batched=False)
    q1, s1 = prop.run(mode, rows, batched=True)
    N = wl.basis.n_states
    tol = 1e-9 if mode == L.MQCT_PROP_AT_RK4 else 1e-6
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    assert d <= tol, f"batched vs plain amplitude diff {d}"
    if mode == L.MQCT_PROP_AT_RK4:
        assert np.array_equal(s0["n_steps"], s1["n_steps"])
    else:
        # the same trajectories (if any) exhaust MAXSTP in both forms
        assert np.array_equal(s0["n_steps"] >= 10000, s1["n_steps"] >= 10000)
    assert np.all(s1["norm_err"] < 1e-6)


def test_at_batched_fp32_coefficients(dev):
    """The batched form with the knot matrices stored in FP32 (batched=2):
    a fixed perturbation of the coupling of ~6e-8 relative, so amplitudes
    stay within 1e-6 of the FP64-coefficient form and the norm is
    conserved as well as there."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=120)
    bb, E = W.at_batch(32, seed=9)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, eps=1e-8, hmin=20.0, at_paths=[i])
    q64, s64 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=1)
    q32, s32 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=2)
    N = wl.basis.n_states
    d = np.max(np.abs(q32[:, :2 * N] - q64[:, :2 * N]))
    assert 0.0 < d <= 1e-6, d
    assert np.array_equal(s32["n_steps"], s64["n_steps"])
    assert np.all(s32["norm_err"] < 1e-6)


N_SHAPES = 23


@pytest.mark.parametrize("shape", list(range(N_SHAPES)))
def test_at_batched_every_gemm_tile(dev, shape):
    """Every tile shape the interval-GEMM tuner may choose gives the plain
    kernel's RK4 trajectories (asymmetric top + atom, j <= 3, 40
    trajectories of uneven length)."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=120)
    bb, E = W.at_batch(40, seed=11)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i])
    q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
    try:
        q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=1)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    assert d <= 1e-9, f"shape {shape}: batched vs plain amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


@pytest.mark.parametrize("shape", [0, 5, 8, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21])
def test_at_batched_fp32_knots_every_layout(dev, shape):
    """FP32 knot matrices through the LDS-staged, the per-wave and the
    fragment-ordered GEMM forms: the same trajectories as the FP64
    coefficients to the FP32 rounding of the coupling (1e-6)."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=120)
    bb, E = W.at_batch(24, seed=13)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i])
    q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
    try:
        q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=2)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    assert 0.0 < d <= 1e-6, f"shape {shape}: fp32-knot batched vs plain amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


@pytest.mark.parametrize("shape,knots", [(0, 1), (12, 1), (15, 1), (16, 1), (15, 2), (16, 2), (17, 1), (18, 1),
                                         (17, 2), (18, 2), (19, 1), (19, 2), (20, 1), (20, 2), (21, 1),
                                         (21, 2)])
def test_at_batched_large_blocks(dev, shape, knots):
    """m-blocks of up to 81 states (asymmetric top j <= 8, 969 states): the
    forms whose waves own several 16-row tiles of a block, and the FP32
    knot layout with its wider fragments, against the plain kernel's RK4
    trajectories (1e-9 with FP64 coefficients, 1e-6 with FP32)."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=8, n_beta=20, n_gamma=40, n_r=120)
    bb, E = W.at_batch(24, seed=17)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i])
    q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
    try:
        q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=knots)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    tol = 1e-9 if knots == 1 else 1e-6
    assert d <= tol, f"shape {shape} knots {knots}: batched vs plain amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


@pytest.mark.parametrize("shape,knots", [(16, 2), (18, 1), (18, 2), (19, 2), (21, 2)])
def test_at_batched_full_blocks(dev, shape, knots):
    """The benchmark's basis (asymmetric top j <= 10 + atom, 1771 states,
    m-blocks of up to 121 states: all eight 16-row tiles of the pipelined
    forms in use), 16 trajectories: the pipelined forms the tuner picks for
    this basis against the plain kernel's RK4 trajectories (1e-9 with FP64
    coefficients, 1e-6 with FP32)."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(n_r=120)
    assert wl.basis.n_states == 1771
    bb, E = W.at_batch(16, seed=29)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i])
    q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
    try:
        q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=knots)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    tol = 1e-9 if knots == 1 else 1e-6
    assert d <= tol, f"shape {shape} knots {knots}: batched vs plain amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


@pytest.mark.parametrize("threads", [512, 1024])
@pytest.mark.parametrize("mode,time_step,hmin", [(L.MQCT_PROP_AT_RK4, 10.0, 20.0),
                                                  (L.MQCT_PROP_AT_ODEINT, 10.0, 20.0),
                                                  (L.MQCT_PROP_AT_ODEINT, 60.0, 10.0)])
def test_at_batched_step_threads(dev, mode, time_step, hmin, threads):
    """The step kernel with 512 threads per trajectory (two states per
    thread fewer, block reductions over eight waves) and with 1024 (two
    states per thread, the earlier k slots loaded with the stage's other
    loads, wave 0 running the next stage's path and bucket chain ahead of
    the states) against the plain kernel, on 969 states where the automatic
    choice would take 256.  At TIME_STEP 60 and hmin 10 rkqs's error
    control decides the steps (the chain then waits for the error at a
    step's last stage)."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=8, n_beta=20, n_gamma=40, n_r=120)
    bb, E = W.at_batch(16, seed=19)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=time_step, eps=1e-8, hmin=hmin,
                                  at_paths=[i])
    q0, s0 = prop.run(mode, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_step_threads(dev.h, threads), "set_step_threads")
    try:
        q1, s1 = prop.run(mode, rows, batched=1)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_step_threads(dev.h, 0), "set_step_threads")
    rej = (s1["n_rhs"] - 6 * s1["n_steps"]) / 5
    print("steps", s1["n_steps"].tolist(), "plain", s0["n_steps"].tolist(), "rejected trials",
          rej.astype(int).tolist(), flush=True)
    if time_step > hmin and mode == L.MQCT_PROP_AT_ODEINT:
        assert np.any(rej > 0)  # rkqs's error control runs
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    tol = 1e-9 if mode == L.MQCT_PROP_AT_RK4 else 1e-6
    assert d <= tol, f"{threads}-thread step kernel vs plain amplitude diff {d}"
    if mode == L.MQCT_PROP_AT_RK4:
        assert np.array_equal(s0["n_steps"], s1["n_steps"])


def test_at_batched_tile_levels(dev):
    """Shapes set per level are used as set (read back after the run, the
    tuner leaving them alone) and give the plain kernel's RK4 trajectories;
    an unknown shape is refused."""
    import ctypes
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(jmax=3, n_beta=12, n_gamma=24, n_r=120)
    bb, E = W.at_batch(24, seed=31)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i])
    q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    want = [8, 11, 19, 18, 21]  # levels 0..4 (up to 31 running trajectories)
    arr = (ctypes.c_int * len(want))(*want)
    L.check(dev.lib.mqct_hip_batched_set_tile_levels(dev.h, 0, len(want), arr), "set_tile_levels")
    bad = (ctypes.c_int * 1)(99)
    assert dev.lib.mqct_hip_batched_set_tile_levels(dev.h, 0, 1, bad) != 0
    q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=1)
    got = (ctypes.c_int * len(want))()
    L.check(dev.lib.mqct_hip_batched_tile_levels(dev.h, 0, len(want), got), "tile_levels")
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")  # back to timing every level
    assert list(got) == want
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    assert d <= 1e-9, f"batched with set levels vs plain amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


@pytest.mark.parametrize("shape", [15, 16, 17, 18, 19, 20, 21, -1])
def test_at_batched_fp32_products(dev, shape):
    """FP32 products (batched=3): the pipelined GEMM forms multiply the FP32
    knot matrices by the weighted amplitudes rounded to FP32 on f32 MFMA
    with FP32 accumulation; the other forms, the step kernel and the
    amplitudes stay FP64.  Against the plain kernel's RK4 trajectories on
    the benchmark basis (m-blocks of up to 121 states): amplitudes within
    1e-6 (1.2e-7 measured) and the same step counts.  shape -1: the tuner's
    choices."""
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import at_first_step
    wl = W.asym_top_atom(n_r=120)
    bb, E = W.at_batch(16, seed=37)
    paths, l, _ = at_first_step(dev, W.asym_top_atom(jmax=1, n_beta=12, n_gamma=24, n_r=120), bb, E,
                                time_step=20.0, rec_cap=20000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(bb)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i])
    q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
    try:
        q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=3)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    print(f"shape {shape}: fp32-product amplitude diff {d:.3e}")
    assert 0.0 < d <= 1e-6, f"shape {shape}: fp32-product batched vs plain amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


def _smooth_v(w):
    """A smooth anisotropic potential of a few hundred cm^-1 on the grid of
    w (the benchmark's site-site potential needs steps below 1 a.u.)."""
    g = w.grid
    b1 = g.xbm + g.xbl * np.asarray(g.xb)
    b2 = g.xbbm + g.xbbl * np.asarray(g.xbb)
    a2, g1, g2 = np.asarray(g.xaa), np.asarray(g.xg), np.asarray(g.xgg)
    A2, B1, B2, G1, G2 = np.meshgrid(a2, b1, b2, g1, g2, indexing="ij")
    ang = (1 + 0.5 * np.cos(B1) + 0.4 * np.cos(B2) + 0.3 * np.cos(A2) * np.sin(B1) * np.sin(B2)
           + 0.2 * np.cos(2 * G1) + 0.2 * np.cos(G2))
    rad = 300.0 * np.exp(-0.8 * (np.asarray(w.R) - 4.0))
    return (rad[:, None, None, None, None, None] * ang[None]).ravel()


@pytest.fixture(scope="module")
def big_blocks(dev):
    """The two-tops basis of the benchmark (2043 states, m-blocks of up to
    369) on a coarse 192-point grid with the smooth potential, built in FP64,
    with AT paths from a first step on j1, j2 <= 1 for 12 trajectories."""
    from mqct_amd.matrix import MatrixBuilder
    wl = W.two_asym_tops(W.C3_JMAX, W.C3_JMAX, emax=W.C3_EMAX, n_ang_each=(2, 4, 3, 2, 4), n_r=40)
    assert max(np.unique(wl.basis.m12, return_counts=True)[1]) > 256
    small = W.two_asym_tops(1, 1, n_ang_each=(2, 4, 3, 2, 4), n_r=40)
    bb, E = W.two_tops_at_batch(12, seed=41)
    MatrixBuilder(dev, small, vgrid=_smooth_v(small)).build("fp64", copy_out=False)
    ps = Propagator(dev, small.basis, small.R, mass_amu=small.mass_amu, from_build=True)
    k = np.sqrt(ps.mass * 2.0 * E * CM_TO_HARTREE)
    l = np.floor(bb * k)
    rws = []
    for i in range(len(bb)):
        rws += ps.trajectories([l[i]], float(E[i]), 1, time_step=20.0)
    _, _, paths = ps.run_record(rws, 20000)
    return wl, paths, l, E


def _big_propagator(dev, big_blocks):
    from mqct_amd.matrix import MatrixBuilder
    wl, paths, l, E = big_blocks
    MatrixBuilder(dev, wl, vgrid=_smooth_v(wl)).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(len(l)):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=20.0, at_paths=[i])
    return wl, prop, rows


@pytest.mark.parametrize("shape,knots", [(15, 1), (17, 1), (18, 1), (20, 1), (18, 2), (18, 3), (21, 3)])
def test_at_batched_pipelined_blocks_over_128(dev, big_blocks, shape, knots):
    """m-blocks above 128 states (the two-tops basis of the benchmark, 2043
    states, m-blocks of up to 369, on a coarse 192-point grid with a smooth
    potential of a few hundred cm^-1): the
    pipelined forms cut such blocks into row groups of 128 rows that
    contract over the whole block.  Against the LDS-staged batched form
    (shape 0, checked against the plain kernel on the same basis below):
    1e-9 with FP64 coefficients, 1e-6 with FP32 coefficients or products,
    same step counts."""
    wl, prop, rows = _big_propagator(dev, big_blocks)
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, 0), "set_tile")
    try:
        q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=1)
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, shape), "set_tile")
        q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=knots)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    N = wl.basis.n_states
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    tol = 1e-9 if knots == 1 else 1e-6
    print(f"shape {shape} knots {knots}: amplitude diff {d:.2e}")
    assert d <= tol, f"shape {shape} knots {knots}: pipelined vs LDS-staged amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])


def test_plain_kernel_lds_above_64k(dev, big_blocks):
    """The plain per-trajectory kernel on 2043 states keeps 2N complex
    amplitudes in 65 632 bytes of LDS (above the 64 KiB a launch gets
    without asking): against the LDS-staged batched form, 1e-9, same
    step counts."""
    wl, prop, rows = _big_propagator(dev, big_blocks)
    N = wl.basis.n_states
    assert 2 * N * 16 + 256 > 64 * 1024
    L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, 0), "set_tile")
    try:
        q0, s0 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=1)
    finally:
        L.check(dev.lib.mqct_hip_batched_set_tile(dev.h, -1), "set_tile")
    q1, s1 = prop.run(L.MQCT_PROP_AT_RK4, rows, batched=0)
    d = np.max(np.abs(q1[:, :2 * N] - q0[:, :2 * N]))
    print(f"plain vs batched on {N} states: amplitude diff {d:.2e}")
    assert d <= 1e-9, f"plain vs LDS-staged batched amplitude diff {d}"
    assert np.array_equal(s0["n_steps"], s1["n_steps"])
