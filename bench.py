"""Benchmark of the MQCT hot paths on one MI355X.

Headline (`value`): AT-MQCT trajectories per second for the batched
propagation of the asymmetric top + atom system (1771 states, 86 779
coupling pairs, 300 spline knots; 4096 trajectories with b uniform in
[0, 18] bohr and E uniform in [200, 2000] cm^-1; amplitudes complex FP64,
Cash-Karp RK45 with tolerance 1e-8 and MIN_TMSTP 20 a.u.; spline
coefficients stored in FP32 by default).  One step = one propagation of the
whole batch, inputs (matrix splines and recorded classical paths) resident
in HBM.

Reported beside it:
  * the same batch with FP64 coefficients, and the cross-section error of
    the FP32-coefficient form against it;
  * the plain FP64 per-trajectory kernel on a sample of the batch (the
    comparison the metric names);
  * the roofline of the dominant kernel (the interval GEMM) from HIP events
    recorded around each of its launches on its stream;
  * matrix-build times of the asymmetric top + atom system (24 x 48 grid,
    300 R) and of the two-asymmetric-tops system (j1, j2 <= 2, 16^5 grid,
    128 R) in FP64, FP32 and split-BF16, and the cross-section error of the
    mixed-precision builds through the AT propagation on a sample;
  * the CPU baseline: the C oracle's AT-RK45 trajectory on one host core.

    python bench.py [--gpus 1] [--steps 1] [--warmup 0]
"""
import argparse
import ctypes
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
    ap.add_argument("--hmin", type=float, default=20.0,
                    help="MIN_TMSTP (a.u.): the step floor of odeint's error control")
    ap.add_argument("--knots", choices=("fp32", "fp64"), default="fp32",
                    help="storage precision of the spline-coefficient matrices of the timed form")
    ap.add_argument("--plain-sample", type=int, default=512)
    ap.add_argument("--xsec-sample", type=int, default=512)
    ap.add_argument("--no-graph", action="store_true",
                    help="time the steps with per-launch HIP events instead of replaying the captured "
                         "graph (for runs under a kernel tracer)")
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
    form = 2 if args.knots == "fp32" else 1

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
    build = {"asym_top_atom": {"fp64": mb.timing()}}
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    rows = []
    for i in range(n):
        rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, eps=1e-8, hmin=args.hmin,
                                  at_paths=[i])

    # prime: knot matrices, kernels, graph (not a step); without graphs when
    # the steps are event-timed (kernel tracers cannot follow graph launches)
    if args.no_graph:
        L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")
    prop.run(L.MQCT_PROP_AT_ODEINT, rows[:8], batched=form)
    for _ in range(args.warmup):
        prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=form)

    # -- timed steps ----------------------------------------------------------
    if args.no_graph:
        L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")  # resets the counts
    t0 = time.perf_counter()
    for _ in range(args.steps):
        q, st = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=form)
    el = time.perf_counter() - t0
    value = n * args.steps / el
    log("batched AT (%s coefficients): %d traj x %d steps in %.2f s -> %.1f traj/s; rhs mean %.0f max %d; "
        "steps max %d; norm err max %.2e" % (args.knots, n, args.steps, el, value, st["n_rhs"].mean(),
                                             st["n_rhs"].max(), st["n_steps"].max(), st["norm_err"].max()))

    # -- one step with HIP events around every GEMM launch --------------------
    if args.no_graph:
        pel = el
    else:
        L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")
        t0 = time.perf_counter()
        prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=form)
        pel = time.perf_counter() - t0
    la, ms, fl, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    L.check(dev.lib.mqct_hip_batched_kernel_stats(dev.h, ctypes.byref(la), ctypes.byref(ms), ctypes.byref(fl),
                                                  ctypes.byref(by)), "kernel_stats")
    L.check(dev.lib.mqct_hip_batched_profile(dev.h, 0), "batched_profile")
    roof = None
    if la.value > 0 and ms.value > 0:
        sec = ms.value * 1e-3 / la.value
        tflops = fl.value / la.value / sec / 1e12
        gbs = by.value / la.value / sec / 1e9
        f_mfma, f_hbm = tflops / PEAK_FP64_TFLOPS, gbs / PEAK_HBM_GBS
        if f_hbm >= f_mfma:
            roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(f_hbm, 4)}
        else:
            roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": PEAK_FP64_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(f_mfma, 4)}
        roof.update({"traffic": traffic_from_profiles(args.knots),
                     "kernel": "gemm_kernel<%s> (interval GEMM, FP64 MFMA 16x16x4)" %
                               ("float" if form == 2 else "double"),
                     "launches": la.value, "avg_launch_us": round(sec * 1e6, 2),
                     "flops_per_launch": fl.value / la.value, "bytes_per_launch": by.value / la.value,
                     "fp64_mfma_tflops": round(tflops, 3), "mfma_frac": round(f_mfma, 4),
                     "hbm_gbs": round(gbs, 1), "hbm_frac": round(f_hbm, 4),
                     "gemm_share_of_step": round(ms.value / (pel * 1e3), 3)})
        log("gemm: %d launches, avg %.1f us, %.2f TFLOP/s (%.1f%% of FP64 peak), %.0f GB/s (%.1f%% of HBM), "
            "%.0f%% of a %.2f s event-timed step" % (la.value, sec * 1e6, tflops, 100 * f_mfma, gbs,
                                                    100 * f_hbm, 100 * ms.value / (pel * 1e3), pel))

    # -- the FP64-coefficient form and its cross sections -----------------------
    other = None
    if form == 2:
        t0 = time.perf_counter()
        if args.no_graph:
            L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")
        q64, _ = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=1)
        oel = time.perf_counter() - t0
        e, nch = rel_err(xsec_functional(prop, q, l, E), xsec_functional(prop, q64, l, E))
        other = {"traj_per_s": round(n / oel, 3), "xsec_rel_err_fp32_vs_fp64_coefficients": e, "channels": nch,
                 "max_amp_diff": float(np.max(np.abs(q[:, :2 * prop.N] - q64[:, :2 * prop.N])))}
        log("batched AT (fp64 coefficients): %.2f s -> %.1f traj/s; xsec fp32 vs fp64 coefficients %.2e over %d "
            "channels" % (oel, n / oel, e, nch))
    else:
        q64 = q

    # -- plain FP64 per-trajectory kernel on a sample ------------------------
    plain = None
    if not args.no_plain and args.plain_sample > 0:
        ns = min(args.plain_sample, n)
        t0 = time.perf_counter()
        qp, sp = prop.run(L.MQCT_PROP_AT_ODEINT, rows[:ns], batched=0)
        pel2 = time.perf_counter() - t0
        dq = float(np.max(np.abs(qp[:, :2 * prop.N] - q64[:ns, :2 * prop.N])))
        plain = {"traj_per_s": round(ns / pel2, 3), "sample": ns, "max_amp_diff_vs_batched_fp64": dq}
        log("plain AT (%d traj): %.2f s -> %.2f traj/s; max |a_batched - a_plain| %.2e" % (ns, pel2, ns / pel2, dq))

    # -- cross-section error of the mixed-precision builds --------------------
    xsec = None
    if not args.no_xsec and args.xsec_sample > 0:
        ns = min(args.xsec_sample, n)
        s64 = xsec_functional(prop, q64[:ns], l[:ns], E[:ns])
        xsec = {"sample": ns}
        for prec in ("fp32", "bf16x3"):
            mb.build(prec, copy_out=False)
            build["asym_top_atom"][prec] = mb.timing()
            pm = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
            pm.load_paths(paths)
            qm, _ = pm.run(L.MQCT_PROP_AT_ODEINT, rows[:ns], batched=1)
            e, nch = rel_err(xsec_functional(pm, qm, l[:ns], E[:ns]), s64)
            xsec[prec] = e
            xsec["channels"] = nch
            log("xsec of the %s build vs fp64: max rel err %.2e over %d channels" % (prec, e, nch))

    # -- matrix-build times ----------------------------------------------------
    if not args.no_build:
        for prec in ("fp32", "bf16x3"):
            if prec not in build["asym_top_atom"]:
                mb.build(prec, copy_out=False)
                build["asym_top_atom"][prec] = mb.timing()
        w3 = W.two_asym_tops()
        m3 = MatrixBuilder(dev, w3)
        build["two_asym_tops"] = {}
        for prec in ("fp64", "fp32", "bf16x3"):
            m3.build(prec, copy_out=False)
            build["two_asym_tops"][prec] = m3.timing()
        log("matrix build total ms: " + ", ".join(
            "%s %s %.1f" % (w, p, build[w][p]["total_ms"]) for w in build for p in build[w]))
    build = {w: {p: {k: round(float(v), 3) for k, v in t.items()} for p, t in d.items()} for w, d in build.items()}

    # -- CPU baseline: the oracle on one core, one trajectory of the batch ----
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dev, wl, paths, rows, st)

    dev.close()
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "traj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el * 1e3 / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "AT-MQCT batched propagation: asymmetric top + atom, 1771 states, "
                               "86779 pairs, 300 knots, 4096 trajectories (b 0-18 bohr, E 200-2000 cm^-1), "
                               "RK45 tol 1e-8, MIN_TMSTP %g a.u., FP64 amplitudes, %s spline coefficients"
                               % (args.hmin, args.knots.upper()),
                   "n_traj": n, "states": int(wl.basis.n_states), "pairs": int(wl.basis.total_size),
                   "knots": int(len(wl.R)), "form": "batched", "graph": not args.no_graph},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_fp64_coefficients": other,
        "plain_fp64": plain,
        "speedup_vs_plain": round(value / plain["traj_per_s"], 2) if plain else None,
        "matrix_build_ms": build,
        "xsec_rel_err_of_builds_vs_fp64": xsec,
        "first_step_paths_truncated": trunc,
    }
    print(json.dumps(out), flush=True)


def traffic_from_profiles(knots):
    """HBM bytes per GEMM launch measured by rocprofv3 counters (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration) in the committed
    profile summary, if present."""
    p = os.path.join(ROOT, "profiles", "batched_gemm_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get(knots, {}).get("bytes_per_launch")


def cpu_baseline(dev, wl, paths, rows, st):
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
    from mqct_amd.matrix import MatrixBuilder
    i = int(np.argmin(np.abs(st["n_rhs"] - st["n_rhs"].mean())))
    M, D, S = MatrixBuilder(dev, wl).build("fp64", copy_out=True)
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
                      "C oracle gcc -O2, 1 thread" % (i, so["n_rhs"], st["n_rhs"].mean())}


if __name__ == "__main__":
    main()
