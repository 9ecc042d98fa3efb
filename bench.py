"""Benchmark of the MQCT hot paths on one MI355X.

Headline (`value`): AT-MQCT trajectories per second for the batched
propagation of the asymmetric top + atom system (1771 states, 86 779
coupling pairs, 300 spline knots; 4096 trajectories with b uniform in
[0, 18] bohr and E uniform in [200, 2000] cm^-1; amplitudes complex FP64,
Cash-Karp RK45 with tolerance 1e-8).  One step = one propagation of the
whole batch, inputs (matrix splines and recorded classical paths) resident
in HBM.

Reported beside it:
  * the plain FP64 per-trajectory kernel on a sample of the same batch
    (the comparison the metric names);
  * matrix-build times of the asymmetric top + atom system (24 x 48 grid,
    300 R) and of the two-asymmetric-tops system (j1, j2 <= 2, 16^5 grid,
    128 R) in FP64, FP32 and split-BF16 (BF16x3);
  * relative cross-section error of the FP32 / BF16x3 matrices against the
    FP64 build, through the same AT propagation on a sample;
  * roofline of the dominant kernel (the interval GEMM, FP64 MFMA), from
    HIP events recorded around each of its launches on its own stream;
  * the CPU baseline: the C oracle's AT-RK45 trajectory on one host core.

    python bench.py [--gpus 1] [--steps 1] [--warmup 0]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "traj/s + matrix-build time vs plain FP64 GPU at ≤1e-5 rel. cross-section error"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def xsec_functional(prop, q, l, E_cm, j_ini=0):
    """Partial-wave sum over the sample: sigma_n = sum_t P_n(t) (2 l_t + 1)
    pi / k_t^2 / (2 j + 1) in bohr^2 (the INELAST_CALC weights,
    propagator.f:4778-4833, with each trajectory standing for its l)."""
    from mqct_amd.basis import CM_TO_HARTREE
    P = prop.opacities(q)
    k2 = prop.mass * 2.0 * np.asarray(E_cm) * CM_TO_HARTREE
    w = (2.0 * np.asarray(l) + 1.0) * np.pi / k2 / (2 * j_ini + 1)
    return (P * w[:, None]).sum(0)


def rel_err(s, s_ref, floor=1e-6):
    """max relative error over channels with sigma >= floor * max sigma."""
    m = s_ref >= floor * s_ref.max()
    return float(np.max(np.abs(s[m] - s_ref[m]) / s_ref[m])), int(m.sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n-traj", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=512)
    ap.add_argument("--xsec-sample", type=int, default=256)
    ap.add_argument("--hmin", type=float, default=20.0,
                    help="MIN_TMSTP (a.u.): the step floor of odeint's error control")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the steps with per-launch HIP events instead of replaying the captured graph "
                         "(for runs under a kernel tracer)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-xsec", action="store_true")
    ap.add_argument("--no-build", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(trajectories would shard across GPUs with replicated splines; not built).", flush=True)
        sys.exit(2)

    from mqct_amd import _lib as L
    from mqct_amd import workloads as W
    from mqct_amd.matrix import Device, MatrixBuilder
    from mqct_amd.propagate import Propagator, at_first_step

    dev = Device(0)
    n = args.n_traj
    bb, E = W.at_batch(n)

    # -- setup: first AT step (classical paths on the j <= 2 basis) ----------
    t0 = time.time()
    small = W.asym_top_atom(jmax=2)
    paths, l, st0 = at_first_step(dev, small, bb, E, time_step=40.0, rec_cap=6000)
    trunc = int(sum(len(p) >= 6000 for p in paths))
    log("first AT step: %d paths in %.1f s, records max %d, truncated %d, failed %d" %
        (n, time.time() - t0, max(len(p) for p in paths), trunc, int(st0["fail"].sum())))

    # -- matrix build of the propagation system (FP64) ----------------------
    wl = W.asym_top_atom()
    mb = MatrixBuilder(dev, wl)
    mb.build("fp64", copy_out=False)
    build = {"asym_top_atom": {"fp64": mb.timing()["total_ms"]}}
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(n):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, eps=1e-8, hmin=args.hmin,
                                  at_paths=[i])

    # prime: knot matrices, kernels, graph (not a step)
    prop.run(L.MQCT_PROP_AT_ODEINT, rows[:8], batched=True)
    for _ in range(args.warmup):
        prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=True)

    # -- timed steps (graph-captured iterations) -------------------------------
    import ctypes
    if args.no_graph:
        L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        q, st = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=True)
    el = time.perf_counter() - t0
    value = n * args.steps / el
    log("batched AT: %d traj x %d steps in %.2f s -> %.1f traj/s; rhs mean %.0f max %d; norm err max %.2e" %
        (n, args.steps, el, value, st["n_rhs"].mean(), st["n_rhs"].max(), st["norm_err"].max()))

    # -- one more step with HIP events around every GEMM launch ---------------
    roof = None
    if args.no_graph:
        q2, pel = q, el
    else:
        L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")
        t0 = time.perf_counter()
        q2, _ = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=True)
        pel = time.perf_counter() - t0
    la, ms, fl = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    L.check(dev.lib.mqct_hip_batched_kernel_stats(dev.h, ctypes.byref(la), ctypes.byref(ms), ctypes.byref(fl)),
            "kernel_stats")
    L.check(dev.lib.mqct_hip_batched_profile(dev.h, 0), "batched_profile")
    same = bool(np.array_equal(q2, q))
    if la.value > 0 and ms.value > 0:
        avg_ms = ms.value / la.value
        ach = fl.value / la.value / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": traffic_from_profiles(),
                "kernel": "gemm_kernel (interval GEMM, FP64 MFMA 16x16x4)",
                "launches": la.value, "avg_launch_us": round(avg_ms * 1e3, 2),
                "flops_per_launch": fl.value / la.value,
                "gemm_share_of_step": round(ms.value / (pel * 1e3), 3),
                "timed_step_ms": round(pel * 1e3, 1), "same_result_as_graph_steps": same}
        log("gemm: %d launches, avg %.1f us, %.2f TFLOP/s (%.1f%% of FP64 peak), %.0f%% of a %.2f s event-timed step"
            % (la.value, avg_ms * 1e3, ach, 100 * ach / PEAK_FP64_TFLOPS, 100 * ms.value / (pel * 1e3), pel))

    # -- plain FP64 per-trajectory kernel on a sample ------------------------
    plain = None
    if not args.no_plain and args.plain_sample > 0:
        ns = min(args.plain_sample, n)
        t0 = time.perf_counter()
        qp, sp = prop.run(L.MQCT_PROP_AT_ODEINT, rows[:ns], batched=False)
        pel = time.perf_counter() - t0
        dq = float(np.max(np.abs(qp[:, :2 * prop.N] - q[:ns, :2 * prop.N])))
        plain = {"traj_per_s": round(ns / pel, 3), "sample": ns, "max_amp_diff_vs_batched": dq}
        log("plain AT (%d traj): %.2f s -> %.2f traj/s; max |a_batched - a_plain| %.2e" % (ns, pel, ns / pel, dq))

    # -- cross-section error of the mixed-precision builds --------------------
    xsec = None
    if not args.no_xsec and args.xsec_sample > 0:
        ns = min(args.xsec_sample, n)
        s64 = xsec_functional(prop, q[:ns], l[:ns], E[:ns])
        xsec = {"sample": ns}
        for prec in ("fp32", "bf16x3"):
            mb.build(prec, copy_out=False)
            build["asym_top_atom"][prec] = mb.timing()["total_ms"]
            pm = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
            pm.load_paths(paths)
            qm, _ = pm.run(L.MQCT_PROP_AT_ODEINT, rows[:ns], batched=True)
            e, nch = rel_err(xsec_functional(pm, qm, l[:ns], E[:ns]), s64)
            xsec[prec] = e
            xsec["channels"] = nch
            log("xsec %s vs fp64: max rel err %.2e over %d channels" % (prec, e, nch))

    # -- matrix-build times ----------------------------------------------------
    if not args.no_build:
        for prec in ("fp32", "bf16x3"):
            if prec not in build["asym_top_atom"]:
                mb.build(prec, copy_out=False)
                build["asym_top_atom"][prec] = mb.timing()["total_ms"]
        w3 = W.two_asym_tops()
        m3 = MatrixBuilder(dev, w3)
        build["two_asym_tops"] = {}
        for prec in ("fp64", "fp32", "bf16x3"):
            m3.build(prec, copy_out=False)
            build["two_asym_tops"][prec] = m3.timing()["total_ms"]
        log("matrix build ms:", json.dumps(build))

    # -- CPU baseline: the oracle on one core, one trajectory of the batch ----
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dev, mb, wl, paths, rows, st)

    dev.close()
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "traj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el * 1e3 / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "AT-MQCT batched propagation: asymmetric top + atom, 1771 states, "
                               "86779 pairs, 300 knots, 4096 trajectories (b 0-18 bohr, E 200-2000 cm^-1), "
                               "RK45 tol 1e-8, FP64 amplitudes",
                   "n_traj": n, "states": int(wl.basis.n_states), "pairs": int(wl.basis.total_size),
                   "knots": int(len(wl.R)), "form": "batched"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_fp64": plain,
        "speedup_vs_plain": round(value / plain["traj_per_s"], 2) if plain else None,
        "matrix_build_ms": build,
        "xsec_rel_err_vs_fp64": xsec,
        "first_step_paths_truncated": trunc,
    }
    print(json.dumps(out), flush=True)


def traffic_from_profiles():
    """HBM bytes per GEMM launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if
    present."""
    p = os.path.join(ROOT, "profiles", "batched_gemm_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(dev, mb, wl, paths, rows, st):
    """The oracle's AT-ODEINT (C restatement of DERIVS_BK_adia + odeint) on
    one host core for the trajectory of the batch whose right-hand-side count
    is closest to the batch mean."""
    try:
        import oracle.oracle as O
    except Exception as e:  # noqa: BLE001
        log("oracle unavailable:", e)
        return None
    from mqct_amd import _lib as L
    from mqct_amd.basis import AMUTOAU, CM_TO_HARTREE
    i = int(np.argmin(np.abs(st["n_rhs"] - st["n_rhs"].mean())))
    M, D, S = mb.build("fp64", copy_out=True)
    b = wl.basis
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, M, D, wl.mass_amu * AMUTOAU)
    dyn.set_path(paths[i])
    r = rows[i]
    t0 = time.perf_counter()
    qo, so, _ = dyn.dotraject(L.MQCT_PROP_AT_ODEINT, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"],
                              hmin=r["hmin"], R_start=r["R_start"])
    el = time.perf_counter() - t0
    log("cpu oracle: trajectory %d (%d rhs) in %.1f s" % (i, so["n_rhs"], el))
    return {"value": round(1.0 / el, 5), "unit": "traj/s", "cores": 1, "kind": "port",
            "sample": "1 trajectory of the batch (index %d, %d right-hand sides; batch mean %.0f), "
                      "C oracle -O2, 1 thread" % (i, so["n_rhs"], st["n_rhs"].mean())}


if __name__ == "__main__":
    main()
