"""Benchmark of the MQCT hot paths on one MI355X.

Headline (`value`): AT-MQCT trajectories per second for the batched
propagation of the asymmetric top + atom system (1771 states, 86 779
coupling pairs, 300 spline knots; 4096 trajectories with b uniform in
[0, 18] bohr and E uniform in [200, 2000] cm^-1; amplitudes complex FP64,
Cash-Karp RK45 (odeint/rkqs) with EPS_ODEINT 1e-8 and MIN_TMSTP 20 a.u.
at TIME_STEP 10 a.u.: the first trial of every step lies at or below
MIN_TMSTP, so rkqs accepts steps of 20 a.u. without its error control
(suppl_routines.f:1245-1253); MIJ_SHIFT on; by default the spline
coefficients are stored in FP32 and the pipelined interval-GEMM forms form
their products in FP32 on f32 MFMA).  One step = one propagation of the
whole batch, inputs (matrix splines and recorded classical paths) resident
in HBM.

Reported beside it:
  * the same batch with FP64 coefficients, and the cross-section error of
    the timed form against it (the metric admits 1e-5 relative);
  * the error-controlled form on a sample of the batch: TIME_STEP 60 a.u.
    puts every first trial above MIN_TMSTP, so rkqs's error control decides
    each step (one rejected trial per step at EPS_ODEINT 1e-8); its rate and
    the cross-section difference of the timed form from it;
  * the plain FP64 per-trajectory kernel on a sample of the batch (the
    comparison the metric names);
  * the roofline of the dominant kernel (the interval GEMM) from HIP events
    recorded around each of its launches on its stream;
  * matrix-build times of the asymmetric top + atom system (24 x 48 grid,
    300 R) and of the two-asymmetric-tops system (16^5 grid, 128 R) in
    FP64, FP32 and split-BF16, and the cross-section error of the
    mixed-precision builds through the AT propagation on a sample;
  * the CPU baseline: the C oracle's AT-RK45 trajectory on one host core.

    python bench.py [--gpus 1] [--steps 1] [--warmup 1]
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
PEAK_FP32_TFLOPS = 157.3
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
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed full-batch runs first (the GEMM tile tuner chooses its shapes in the first)")
    ap.add_argument("--n-traj", type=int, default=4096)
    ap.add_argument("--hmin", type=float, default=20.0,
                    help="MIN_TMSTP (a.u.): the step floor of odeint's error control")
    ap.add_argument("--time-step", type=float, default=10.0,
                    help="TIME_STEP (a.u.) of the timed form: odeint's first trial step is TIME_STEP/sqrt(U/300)")
    ap.add_argument("--ctrl-sample", type=int, default=1024,
                    help="trajectories of the error-controlled run (TIME_STEP 60: every first trial above "
                         "MIN_TMSTP); 0 skips it")
    ap.add_argument("--knots", choices=("fp32", "fp64", "fp32p"), default="fp32p",
                    help="storage precision of the spline-coefficient matrices of the timed form "
                         "(fp32p: FP32 storage and FP32 MFMA products in the pipelined GEMM forms)")
    ap.add_argument("--step-threads", type=int, default=0, choices=(0, 256, 512, 1024),
                    help="threads per trajectory of the batched step kernel (0: the library's choice)")
    ap.add_argument("--tile-levels", default="",
                    help="comma-separated interval-GEMM shape per level floor(log2(running trajectories)), "
                         "from level 0 up (-1: chosen by timing), for the timed form's knot precision: "
                         "repeats a measured run's choices in a profiled one")
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--xsec-sample", type=int, default=512)
    ap.add_argument("--no-graph", action="store_true",
                    help="time the steps with per-launch HIP events instead of replaying the captured "
                         "graph (for runs under a kernel tracer)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the run with the other coefficient precision")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-xsec", action="store_true")
    ap.add_argument("--no-build", action="store_true")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the two-asymmetric-tops forms (builds, surrogate, trajectory sample)")
    ap.add_argument("--c3-traj", type=int, default=16,
                    help="two-tops AT trajectories (a sample of the 1024-trajectory batch)")
    ap.add_argument("--exp-terms", type=int, default=1,
                    help="also build the two-tops matrix through the PES expansion (0: skip)")
    ap.add_argument("--c3-plain-steps", type=int, default=40,
                    help="steps of the bounded plain-kernel run that gives its right-hand-side rate")
    ap.add_argument("--c3-dt", type=float, default=0.5,
                    help="TIME_STEP (a.u.) of the two-tops rk4 propagation")
    ap.add_argument("--c3-dt2", type=float, default=0.0,
                    help="also propagate the FP64 two-tops build at this TIME_STEP and report the cross-section "
                         "change (a step-size convergence check; 0 skips it)")
    ap.add_argument("--c3-chunk", type=int, default=1024,
                    help="two-tops trajectories per propagate call (each call is logged when it ends)")
    ap.add_argument("--c3-only", action="store_true",
                    help="run only the two-asymmetric-tops section (no headline line)")
    ap.add_argument("--c3-dt-study", type=str, default="",
                    help="comma-separated integrator settings: a number is an rk4 TIME_STEP (a.u.), e<eps> is "
                         "odeint with that EPS_ODEINT (MIN_TMSTP --c3-hmin); only propagate the FP64 two-tops "
                         "build's sample (c3-traj trajectories) with each and report the cross-section change "
                         "between consecutive ones (no headline line)")
    ap.add_argument("--c3-hmin", type=float, default=0.25,
                    help="MIN_TMSTP (a.u.) of the two-tops odeint runs of the integrator study")
    ap.add_argument("--sur-samples", type=int, default=4_000_000)
    ap.add_argument("--sur-steps", type=int, default=30000)
    ap.add_argument("--sur-batch", type=int, default=8192)
    ap.add_argument("--sur-rmin", type=float, default=4.0,
                    help="the surrogate covers R >= this (bohr); the plain PES fills the grid below")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (4 launches, each reading and writing "
                         "256 MiB) for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(trajectories would shard across GPUs with replicated splines; not built).", flush=True)
        sys.exit(2)

    from mqct_amd import _lib as L
    from mqct_amd import workloads as W
    from mqct_amd.matrix import Device, MatrixBuilder
    from mqct_amd.propagate import Propagator, at_first_step

    dev = Device(0)
    if args.c3_dt_study:
        rec = c3_dt_study(dev, args, args.c3_dt_study.split(","))
        dev.close()
        print(json.dumps({"two_asym_tops_dt_study": rec}), flush=True)
        return 0
    if args.c3_only:
        c3 = two_tops_section(dev, args)
        dev.close()
        print(json.dumps({"two_asym_tops": c3}), flush=True)
        return 0
    n = args.n_traj
    bb, E = W.at_batch(n)
    form = {"fp64": 1, "fp32": 2, "fp32p": 3}[args.knots]

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
    L.check(dev.lib.mqct_hip_batched_set_step_threads(dev.h, args.step_threads), "set_step_threads")
    if args.tile_levels:
        lv_set = [int(x) for x in args.tile_levels.split(",")]
        arr = (ctypes.c_int * len(lv_set))(*lv_set)
        L.check(dev.lib.mqct_hip_batched_set_tile_levels(dev.h, form - 1, len(lv_set), arr),
                "set_tile_levels")
    def batch_rows(time_step, count=n):
        out = []
        for i in range(count):
            out += prop.trajectories([l[i]], float(E[i]), 1, time_step=time_step, eps=1e-8, hmin=args.hmin,
                                     at_paths=[i])
        return out
    rows = batch_rows(args.time_step)

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
    # each rkqs trial after the first of a step costs 5 more right-hand sides
    rejected = (st["n_rhs"] - 6 * st["n_steps"]) / 5.0
    integ = {"rhs_mean": float(st["n_rhs"].mean()), "rhs_max": int(st["n_rhs"].max()),
             "steps_mean": float(st["n_steps"].mean()), "rejected_trials_mean": float(rejected.mean()),
             "trajectories_with_rejections": int((rejected > 0).sum()), "norm_err_max": float(st["norm_err"].max()),
             "maxstp_reached": int((st["n_steps"] >= 10000).sum())}
    log("batched AT (%s coefficients): %d traj x %d steps in %.2f s -> %.1f traj/s; rhs mean %.0f max %d; "
        "steps mean %.0f, rejected trials mean %.0f; norm err max %.2e" % (
            args.knots, n, args.steps, el, value, st["n_rhs"].mean(), st["n_rhs"].max(), st["n_steps"].mean(),
            rejected.mean(), st["norm_err"].max()))

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
    # GEMM and step time per level floor(log2(running trajectories))
    ntr = ctypes.c_int()
    L.check(dev.lib.mqct_hip_batched_prof_trace(dev.h, 0, None, ctypes.byref(ntr)), "prof_trace")
    trace = np.zeros(5 * max(ntr.value, 1))
    L.check(dev.lib.mqct_hip_batched_prof_trace(dev.h, ntr.value, L.dptr(trace), ctypes.byref(ntr)), "prof_trace")
    levels = {}
    for a0, nit, gms, _sh, sms in trace[:5 * ntr.value].reshape(-1, 5):
        key = "2^%d" % int(np.log2(max(a0, 1)))
        e = levels.setdefault(key, {"iterations": 0, "gemm_ms": 0.0, "step_ms": 0.0})
        e["iterations"] += int(nit)
        e["gemm_ms"] += float(gms)
        e["step_ms"] += float(sms)
    levels = {k: {"iterations": v["iterations"], "gemm_ms": round(v["gemm_ms"], 1), "step_ms": round(v["step_ms"], 1),
                  "gemm_us_per_it": round(1e3 * v["gemm_ms"] / max(v["iterations"], 1), 1),
                  "step_us_per_it": round(1e3 * v["step_ms"] / max(v["iterations"], 1), 1)}
              for k, v in sorted(levels.items(), key=lambda kv: -int(kv[0][2:]))}
    log("per level: " + "; ".join("%s %d it, gemm %.1f us, step %.1f us" % (k, v["iterations"], v["gemm_us_per_it"],
                                                                             v["step_us_per_it"])
                                   for k, v in levels.items()))
    L.check(dev.lib.mqct_hip_batched_profile(dev.h, 0), "batched_profile")
    FORMS = ("lds", "direct", "stream", "queue", "lds_queue", "pipe", "pipe_dyn", "stream_list")

    def shape_desc(sh):
        fo, ge, gk, grt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.check(dev.lib.mqct_hip_batched_tile(dev.h, sh, ctypes.byref(fo), ctypes.byref(ge), ctypes.byref(gk),
                                              ctypes.byref(grt)), "tile")
        return "%s e%d k%d r%d" % (FORMS[fo.value], ge.value, gk.value, grt.value)
    lv = (ctypes.c_int * 24)()
    L.check(dev.lib.mqct_hip_batched_tile_levels(dev.h, form - 1, 24, lv), "tile_levels")
    tile = {"first": shape_desc(-1),
            "per_level": {"2^%d" % i: shape_desc(v) for i, v in enumerate(lv) if v >= 0}}
    roof = None
    if la.value > 0 and ms.value > 0:
        sec = ms.value * 1e-3 / la.value
        tflops = fl.value / la.value / sec / 1e12
        gbs = by.value / la.value / sec / 1e9
        # the pipelined GEMM forms (the full and middle levels) run f32 MFMA
        # with FP32 products, f64 MFMA otherwise
        peak_mfma = PEAK_FP32_TFLOPS if form == 3 else PEAK_FP64_TFLOPS
        f_mfma, f_hbm = tflops / peak_mfma, gbs / PEAK_HBM_GBS
        if f_hbm >= f_mfma:
            roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(f_hbm, 4)}
        else:
            roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": peak_mfma,
                    "unit": "TFLOP/s", "frac": round(f_mfma, 4)}
        roof.update({"traffic": traffic_from_profiles(args.knots),
                     "kernel": "interval GEMM over %s knot matrices, %s (gemm_*_kernel: form and tile shape "
                               "chosen per running-trajectory level)"
                               % ("FP64" if form == 1 else "FP32",
                                  "f32 MFMA 16x16x4 with FP32 products in the pipelined forms, f64 MFMA 16x16x4 "
                                  "in the others" if form == 3 else "f64 MFMA 16x16x4"),
                     "mfma_peak_tflops": peak_mfma,
                     "tile": tile, "launches": la.value, "avg_launch_us": round(sec * 1e6, 2),
                     "flops_per_launch": fl.value / la.value, "bytes_per_launch": by.value / la.value,
                     "mfma_tflops": round(tflops, 3), "mfma_frac": round(f_mfma, 4),
                     "hbm_gbs": round(gbs, 1), "hbm_frac": round(f_hbm, 4),
                     "gemm_share_of_step": round(ms.value / (pel * 1e3), 3), "levels": levels})
        log("gemm: %d launches, avg %.1f us, %.2f TFLOP/s (%.1f%% of the %.1f TFLOP/s MFMA peak), %.0f GB/s "
            "(%.1f%% of HBM), %.0f%% of a %.2f s event-timed step"
            % (la.value, sec * 1e6, tflops, 100 * f_mfma, peak_mfma, gbs, 100 * f_hbm,
               100 * ms.value / (pel * 1e3), pel))

    # -- the other coefficient precision and the cross-section difference -----
    other = None
    if not args.no_other:
        oform = 2 if form == 1 else 1
        t0 = time.perf_counter()
        if args.no_graph:
            L.check(dev.lib.mqct_hip_batched_profile(dev.h, 1), "batched_profile")
        qo, _ = prop.run(L.MQCT_PROP_AT_ODEINT, rows, batched=oform)
        oel = time.perf_counter() - t0
        # the reduced-precision form against the FP64 coefficients
        qx, q64 = (qo, q) if form == 1 else (q, qo)
        xname = "fp32" if form == 1 else args.knots
        e, nch = rel_err(xsec_functional(prop, qx, l, E), xsec_functional(prop, q64, l, E))
        other = {"knots": "fp32" if oform == 2 else "fp64", "traj_per_s": round(n / oel, 3),
                 "xsec_rel_err_vs_fp64_coefficients": e, "compared_form": xname, "channels": nch,
                 "max_amp_diff": float(np.max(np.abs(qx[:, :2 * prop.N] - q64[:, :2 * prop.N])))}
        log("batched AT (%s coefficients): %.2f s -> %.1f traj/s; xsec %s vs fp64 coefficients %.2e over %d "
            "channels" % (other["knots"], oel, n / oel, xname, e, nch))
    else:
        q64 = q

    # -- the error-controlled form on a sample (TIME_STEP 60) -------------------
    ctrl = None
    if args.ctrl_sample > 0:
        nc = min(args.ctrl_sample, n)
        rows_c = batch_rows(60.0, nc)
        t0 = time.perf_counter()
        qc, sc = prop.run(L.MQCT_PROP_AT_ODEINT, rows_c, batched=form)
        cel = time.perf_counter() - t0
        ec, nchc = rel_err(xsec_functional(prop, q[:nc], l[:nc], E[:nc]), xsec_functional(prop, qc, l[:nc], E[:nc]))
        ctrl = {"traj_per_s": round(nc / cel, 3), "n_traj": nc, "time_step": 60.0,
                "rhs_mean": float(sc["n_rhs"].mean()), "steps_mean": float(sc["n_steps"].mean()),
                "rejected_trials_mean": float(((sc["n_rhs"] - 6 * sc["n_steps"]) / 5.0).mean()),
                "maxstp_reached": int((sc["n_steps"] >= 10000).sum()),
                "xsec_rel_err_of_timed_form": ec, "channels": nchc}
        log("error-controlled (TIME_STEP 60), %d traj: %.2f s -> %.1f traj/s; rhs mean %.0f, rejected trials mean "
            "%.0f; timed form's xsec vs it %.2e" % (nc, cel, nc / cel, sc["n_rhs"].mean(),
                                                    ctrl["rejected_trials_mean"], ec))

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

    # -- matrix-build times: a warm build of each form (the first build of a
    #    basis and grid also forms the basis tables, and the first of each
    #    form converts the weights and times the contraction variants) ------
    if not args.no_build:
        for prec in ("fp64", "fp32", "bf16x3"):
            if prec not in build["asym_top_atom"]:
                mb.build(prec, copy_out=False)
            mb.build(prec, copy_out=False)
            build["asym_top_atom"][prec] = mb.timing()
        log("matrix build total ms: " + ", ".join(
            "%s %s %.1f" % (w, p, build[w][p]["total_ms"]) for w in build for p in build[w]))
    build = {w: {p: {k: round(float(v), 3) for k, v in t.items()} for p, t in d.items()} for w, d in build.items()}

    # -- two asymmetric tops: precision and surrogate forms through 256 AT
    #    trajectories -------------------------------------------------------
    c3 = None
    if not args.no_c3:
        c3 = two_tops_section(dev, args)
        build["two_asym_tops"] = {p: c3[p]["build_ms"] for p in ("fp64", "fp32", "bf16x3")}
        if "expansion" in c3:
            build["two_asym_tops"]["expansion"] = c3["expansion"]["build_ms"]

    # -- CPU baseline: the oracle on one core, one trajectory of the batch ----
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dev, wl, paths, rows, st)

    dev.close()
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "traj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el * 1e3 / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64/f32" if form == 3 else "f64", "data": "synthetic",
        "config": {"workload": "AT-MQCT batched propagation: asymmetric top + atom, 1771 states, "
                               "86779 pairs, 300 knots, 4096 trajectories (b 0-18 bohr, E 200-2000 cm^-1), "
                               "Cash-Karp RK45 (odeint/rkqs) EPS_ODEINT 1e-8, TIME_STEP %g a.u., MIN_TMSTP "
                               "%g a.u. (first trials at or below MIN_TMSTP: steps accepted at the floor "
                               "without error control), MIJ_SHIFT, FP64 amplitudes, %s"
                               % (args.time_step, args.hmin, {"fp64": "FP64 spline coefficients",
                                              "fp32": "FP32 spline coefficients, FP64 products",
                                              "fp32p": "FP32 spline coefficients, FP32 products in the "
                                                       "pipelined GEMM forms"}[args.knots]),
                   "n_traj": n, "states": int(wl.basis.n_states), "pairs": int(wl.basis.total_size),
                   "knots": int(len(wl.R)), "form": "batched", "graph": not args.no_graph},
        "roofline": roof,
        "cpu_baseline": cpu,
        "integrator": integ,
        "error_controlled_sample": ctrl,
        "batched_other_coefficients": other,
        "plain_fp64": plain,
        "speedup_vs_plain": round(value / plain["traj_per_s"], 2) if plain else None,
        "matrix_build_ms": build,
        "xsec_rel_err_of_builds_vs_fp64": xsec,
        "two_asym_tops": c3,
        "first_step_paths_truncated": trunc,
    }
    print(json.dumps(out), flush=True)


def copy_calibration():
    """A plain elementwise device copy (256 MiB read + 256 MiB written per
    launch, 4 launches): the counter reading per launch divided into
    262144 KiB is the calibration factor of FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = (256 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(4):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print("copy calibration: 4 launches of %d bytes read + %d bytes written" % (4 * n, 4 * n), flush=True)
    return 0


def elem_err(M, M64):
    """Matrix-element error against the FP64 build, relative to the largest
    |M| at the same R (elements that cancel to almost nothing carry only
    absolute accuracy): max and 99.9th percentile over elements and R."""
    scale = np.maximum(np.abs(M64).max(axis=0, keepdims=True), 1e-300)
    e = np.abs(M - M64) / scale
    return {"max": float(e.max()), "p999": float(np.percentile(e, 99.9))}


def two_tops_section(dev, args):
    """The two-asymmetric-tops system of the metric (the CHNLS_LIST of the
    pair levels with j1, j2 <= 4 and E1 + E2 <= 236.5 cm^-1: 2043 states,
    282 179 pairs; 16^5 grid, 128 R): FP64, FP32 and split-BF16 builds with
    the plain PES, and a split-BF16 build on the surrogate PES grid; the
    FP64, split-BF16 and surrogate matrices drive the same sample of AT
    trajectories (paths from a first
    step on j1, j2 <= 1), batched with FP64 coefficients, rk4 at dt = c3_dt
    a.u. (0.5: norm drift below 1e-6), and the cross sections are compared
    with the FP64 build's.  The plain FP64 kernel's rate is measured on a
    bounded run (its first c3_plain_steps steps of up to 64 trajectories, one
    workgroup each) and
    scaled by the sample's mean number of right-hand sides per trajectory."""
    from mqct_amd import _lib as L
    from mqct_amd import workloads as W
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import Propagator, at_first_step
    from mqct_amd.surrogate import train_surrogate
    n = args.c3_traj
    # long propagations report the trajectories still running every minute
    L.check(dev.lib.mqct_hip_set_progress(dev.h, 60.0), "set_progress")
    wl = W.two_asym_tops_c3()
    bb, E = W.two_tops_at_batch(1024)
    sel = np.linspace(0, 1023, n).astype(int) if n < 1024 else np.arange(1024)
    bb, E = bb[sel], E[sel]
    t0 = time.time()
    paths, l, st0 = at_first_step(dev, W.two_asym_tops(jmax1=1, jmax2=1), bb, E, time_step=40.0, rec_cap=8000)
    log("two tops: first AT step %d paths in %.1f s, records max %d, failed %d" %
        (n, time.time() - t0, max(len(p) for p in paths), int(st0["fail"].sum())))
    out = {"states": int(wl.basis.n_states), "pairs": int(wl.basis.total_size), "channels": int(wl.basis.n_channels),
           "grid_points": int(wl.grid.G), "r_points": int(len(wl.R)), "n_traj": n,
           "sample_of": 1024, "integrator": "AT rk4, dt %g a.u." % args.c3_dt}

    def rows_for(prop, count=n, max_steps=3000000):
        rows = []
        for i in range(count):
            rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i], max_steps=max_steps)
            rows[-1]["dt"] = args.c3_dt
        return rows

    def propagate(batched=1, dt=None, tag=""):
        prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
        prop.load_paths(paths)
        rows = rows_for(prop)
        if dt is not None:
            for r in rows:
                r["dt"] = dt
        qs, sts, el = [], [], 0.0
        for c0 in range(0, n, args.c3_chunk):
            t0 = time.perf_counter()
            q, st = prop.run(L.MQCT_PROP_AT_RK4, rows[c0:c0 + args.c3_chunk], batched=batched)
            el += time.perf_counter() - t0
            qs.append(q)
            sts.append(st)
            if n > args.c3_chunk:
                log("two tops %s: trajectories %d-%d done, %.1f s so far" % (tag, c0, c0 + len(q) - 1, el))
        q = np.concatenate(qs)
        st = {k: np.concatenate([s[k] for s in sts]) for k in sts[0]}
        return prop, q, st, n / el

    mb = MatrixBuilder(dev, wl)
    M64 = None
    sig64 = None
    for prec in ("fp64", "fp32", "bf16x3"):
        M, D, S = mb.build(prec)
        tm = {k: round(float(v), 3) for k, v in mb.timing().items()}
        rec = {"build_ms": tm}
        if prec == "fp32":
            # build time and element error only (the sample propagates the
            # FP64 and split-BF16 builds)
            rec["elem_err"] = elem_err(M, M64)
            rec["build_speedup_vs_fp64"] = round(out["fp64"]["build_ms"]["total_ms"] / tm["total_ms"], 3)
            out[prec] = rec
            log("two tops fp32: build %.1f ms (contract %.1f), elem err max %.2e" % (
                tm["total_ms"], tm["contract_ms"], rec["elem_err"]["max"]))
            continue
        prop, q, st, tps = propagate(1, tag=prec)
        sig = xsec_functional(prop, q, l, E)
        rec.update({"batched_traj_per_s": round(tps, 3), "norm_err_max": float(st["norm_err"].max()),
                    "rhs_mean": float(st["n_rhs"].mean())})
        if prec == "fp64":
            M64, sig64 = M, sig
            if args.c3_dt2 > 0:
                p2, q2, s2, tps2 = propagate(1, dt=args.c3_dt2, tag="fp64 dt %g" % args.c3_dt2)
                e2, nch2 = rel_err(xsec_functional(p2, q2, l, E), sig)
                rec["dt_study"] = {"dt": args.c3_dt2, "traj_per_s": round(tps2, 3), "xsec_rel_change": e2,
                                   "channels": nch2, "norm_err_max": float(s2["norm_err"].max())}
                log("two tops fp64 at dt %g: %.3f traj/s, cross sections differ by %.2e from dt %g" % (
                    args.c3_dt2, tps2, e2, args.c3_dt))
            # the plain kernel on a bounded run: right-hand sides per second
            pr = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
            pr.load_paths(paths)
            rows_b = rows_for(pr, min(64, n), max_steps=args.c3_plain_steps)
            t0 = time.perf_counter()
            _, sp = pr.run(L.MQCT_PROP_AT_RK4, rows_b, batched=0)
            pel = time.perf_counter() - t0
            rhs_rate = float(sp["n_rhs"].sum()) / pel
            rec["plain_rhs_per_s"] = round(rhs_rate, 1)
            rec["plain_traj_per_s_est"] = round(rhs_rate / st["n_rhs"].mean(), 5)
            rec["plain_note"] = ("plain FP64 kernel: %d trajectories x first %d steps in %.1f s, scaled by the "
                                 "sample's mean right-hand sides per trajectory" % (len(rows_b),
                                                                                    args.c3_plain_steps, pel))
            rec["speedup_batched_vs_plain"] = round(tps / rec["plain_traj_per_s_est"], 2)
        else:
            rec["elem_err"] = elem_err(M, M64)
            rec["xsec_rel_err"], rec["channels"] = rel_err(sig, sig64)
            rec["build_speedup_vs_fp64"] = round(out["fp64"]["build_ms"]["total_ms"] / tm["total_ms"], 3)
        out[prec] = rec
        log("two tops %s: build %.1f ms (contract %.1f), batched %.3f traj/s%s" % (
            prec, tm["total_ms"], tm["contract_ms"], tps,
            (", plain est %.4f traj/s" % rec["plain_traj_per_s_est"]) if prec == "fp64" else
            ", xsec err %.2e, elem err max %.2e" % (rec["xsec_rel_err"], rec["elem_err"]["max"])))
    if args.exp_terms:
        out["expansion"] = expansion_section(dev, mb, wl, M64, out["fp64"]["build_ms"]["total_ms"])
    if args.sur_steps <= 0:
        return out
    # the learned potential: trained on the GPU, fills the V grid, split-BF16 build
    sur, stats = train_surrogate(dev, wl, n_samples=args.sur_samples, steps=args.sur_steps, batch=args.sur_batch,
                                 r_lo=args.sur_rmin, log=log)
    mb2 = MatrixBuilder(dev, wl)  # plain V grid: the fallback below r_min
    par = L.as_f64(wl.pes_par)
    R = L.as_f64(wl.R)
    t0 = time.perf_counter()
    L.check(dev.lib.mqct_hip_eval_pes(dev.h, len(R), L.dptr(R), wl.pes_kind, len(par), L.dptr(par)), "eval_pes")
    plain_fill_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sur.fill_vgrid(wl.R)
    fill_s = time.perf_counter() - t0
    M, D, S = mb2.build("bf16x3")
    tm = {k: round(float(v), 3) for k, v in mb2.timing().items()}
    prop, q, st, tps = propagate(1, tag="surrogate")
    sig = xsec_functional(prop, q, l, E)
    e, nch = rel_err(sig, sig64)
    stats.update({"grid_fill_s": round(fill_s, 4), "plain_pes_grid_fill_gpu_s": round(plain_fill_s, 4),
                  "grid_points": int(wl.grid.G) * len(wl.R), "build_ms": tm, "elem_err": elem_err(M, M64),
                  "xsec_rel_err": e, "channels": nch, "batched_traj_per_s": round(tps, 3)})
    out["surrogate_bf16x3"] = stats
    log("two tops surrogate: fit %.1f s, grid fill %.3f s, xsec err %.2e, elem err max %.2e" % (
        stats["fit_s"], fill_s, e, stats["elem_err"]["max"]))
    return out


def c3_dt_study(dev, args, specs):
    """Integrator convergence of the two-tops propagation: the sample of
    the 1024-trajectory batch (c3_traj trajectories, the same selection as
    the two-tops section) on the FP64 build, batched with FP64
    coefficients, with each integrator setting of specs: a number is an
    rk4 TIME_STEP (rk4's global error goes as dt^4, so halving dt leaves
    about 1/16 of it), e<eps> is odeint (Cash-Karp with error control) at
    EPS_ODEINT eps and MIN_TMSTP c3_hmin.  The change between consecutive
    settings is reported as the largest relative change of the sample's
    cross sections over the channels above 1e-6, 1e-4 and 1e-2 of the
    largest, with the norm drift and the right-hand sides per trajectory."""
    from mqct_amd import _lib as L
    from mqct_amd import workloads as W
    from mqct_amd.matrix import MatrixBuilder
    from mqct_amd.propagate import Propagator, at_first_step
    n = args.c3_traj
    L.check(dev.lib.mqct_hip_set_progress(dev.h, 60.0), "set_progress")
    wl = W.two_asym_tops_c3()
    bb, E = W.two_tops_at_batch(1024)
    sel = np.linspace(0, 1023, n).astype(int) if n < 1024 else np.arange(1024)
    bb, E = bb[sel], E[sel]
    paths, l, st0 = at_first_step(dev, W.two_asym_tops(jmax1=1, jmax2=1), bb, E, time_step=40.0, rec_cap=8000)
    MatrixBuilder(dev, wl).build("fp64", copy_out=False)
    prop = Propagator(dev, wl.basis, wl.R, mass_amu=wl.mass_amu, from_build=True)
    prop.load_paths(paths)
    out = {"states": int(wl.basis.n_states), "n_traj": n, "sample_of": 1024, "runs": []}
    prev = None
    for spec in specs:
        ode = spec.startswith("e")
        rows = []
        for i in range(n):
            if ode:
                rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, eps=float(spec[1:]),
                                          hmin=args.c3_hmin, at_paths=[i])
            else:
                rows += prop.trajectories([l[i]], float(E[i]), 1, time_step=10.0, at_paths=[i], max_steps=3000000)
                rows[-1]["dt"] = float(spec)
        t0 = time.perf_counter()
        q, st = prop.run(L.MQCT_PROP_AT_ODEINT if ode else L.MQCT_PROP_AT_RK4, rows, batched=1)
        el = time.perf_counter() - t0
        sig = xsec_functional(prop, q, l, E)
        dt = ("odeint eps %s hmin %g" % (spec[1:], args.c3_hmin)) if ode else float(spec)
        run = {"dt": dt, "traj_per_s": round(n / el, 3), "norm_err_max": float(st["norm_err"].max()),
               "rhs_mean": float(st["n_rhs"].mean()), "rhs_max": int(st["n_rhs"].max())}
        if ode:
            run["maxstp_reached"] = int((st["n_steps"] >= 10000).sum())
        run["sigma"] = [float(x) for x in sig]
        if prev is not None:
            run["xsec_rel_change_vs_dt"] = prev[0]
            for fl in (1e-6, 1e-4, 1e-2):
                e, nch = rel_err(sig, prev[1], floor=fl)
                run["xsec_rel_change_floor_%g" % fl] = e
                run["channels_floor_%g" % fl] = nch
        out["runs"].append(run)
        log("two tops integrator study: %s, %.3f traj/s, rhs mean %.0f, norm drift %.2e%s" % (
            dt, run["traj_per_s"], run["rhs_mean"], run["norm_err_max"],
            "" if prev is None else ", cross sections change from %s by %.2e / %.2e / %.2e (floors 1e-6 / "
            "1e-4 / 1e-2)" % (prev[0], run["xsec_rel_change_floor_1e-06"], run["xsec_rel_change_floor_0.0001"],
                              run["xsec_rel_change_floor_0.01"])))
        prev = (dt, sig)
    return out


def expansion_section(dev, mb, wl, M64, fp64_build_ms):
    """The PES-expansion route to the same two-tops matrix: the 312 terms of
    L1, L2 <= 4 with even NJU1, NJU2 (both partners have a C2 axis) are
    projected from the FP64 V grid on the GPU (CALC_EXPANSION_TERMS), then
    the matrix is built from them (EXPANSION_MATRIX_ELEMENT: angular
    coefficients, one GEMM over the terms, cutoff and spline).  elem_err is
    the expansion's truncation against the quadrature build, not a
    rounding error."""
    from mqct_amd import expansion as X
    terms = X.terms_type0(4, 4, du1=2, du2=2)
    ex = X.Expansion(dev, wl, terms)
    t0 = time.perf_counter()
    ex.project()
    proj_s = time.perf_counter() - t0
    ex.build(copy_out=False)          # first build: CG table and buffers
    M, D, S = ex.build()
    t = {k: round(float(v), 3) for k, v in ex.timing().items()}
    bt = {k: round(float(v), 3) for k, v in mb.timing().items()}
    rec = {"terms": len(terms), "project_wall_s": round(proj_s, 3), "device_ms": t,
           "build_ms": bt, "build_speedup_vs_fp64_quadrature": round(fp64_build_ms / bt["total_ms"], 2),
           "elem_err_vs_quadrature": elem_err(M, M64)}
    log("two tops expansion: %d terms, projection %.1f + %.1f ms, build %.1f ms (angular %.1f, GEMM %.1f) = "
        "%.1fx the FP64 quadrature build; truncation vs quadrature max %.2e" % (
            len(terms), t["project_terms_ms"], t["project_gemm_ms"], bt["total_ms"], t["angular_ms"],
            t["gemm_ms"], rec["build_speedup_vs_fp64_quadrature"], rec["elem_err_vs_quadrature"]["max"]))
    return rec


def traffic_from_profiles(knots):
    """HBM bytes per GEMM launch measured by rocprofv3 counters (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration) in the committed
    profile summary, if present."""
    p = os.path.join(ROOT, "profiles", "batched_gemm_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get(knots, {}).get("bytes_per_launch")


def cpu_baseline(dev, wl, paths, rows, st, n_sample=4):
    """The oracle's AT-ODEINT (C restatement of DERIVS_BK_adia + odeint) on
    one host core for the n_sample trajectories of the batch whose
    right-hand-side counts are closest to the batch mean (about 15 s of CPU
    work); value = trajectories per second over the sample."""
    try:
        import oracle.oracle as O
    except Exception as e:  # noqa: BLE001
        log("oracle unavailable:", e)
        return None
    from mqct_amd import _lib as L
    from mqct_amd.basis import AMUTOAU, CM_TO_HARTREE
    from mqct_amd.matrix import MatrixBuilder
    pick = baseline_sample(st["n_rhs"], n_sample)
    M, D, S = MatrixBuilder(dev, wl).build("fp64", copy_out=True)
    b = wl.basis
    # the GPU propagates M(R) - M(R_max) (MIJ_SHIFT=YES, matrix_ini.f:2226-2230)
    dyn = O.DynOracle(b.j12, b.m12, b.Ej * CM_TO_HARTREE, b.ind_mat, wl.R, O.mij_shift(M), D,
                      wl.mass_amu * AMUTOAU)
    el, nrhs = 0.0, 0
    for i in pick:
        dyn.set_path(paths[i])
        r = rows[i]
        t0 = time.perf_counter()
        qo, so, _ = dyn.dotraject(L.MQCT_PROP_AT_ODEINT, 1, r["l_real"], r["E_coll"], r["dt"], eps=r["eps"],
                                  hmin=r["hmin"], R_start=r["R_start"])
        el += time.perf_counter() - t0
        nrhs += int(so["n_rhs"])
    log("cpu oracle: %d trajectories (%d rhs) in %.1f s" % (len(pick), nrhs, el))
    out = {"value": round(len(pick) / el, 5), "unit": "traj/s", "cores": 1, "kind": "port",
           "sample": "%d trajectories of the batch (indices %s, %.0f right-hand sides each on average; batch "
                     "mean %.0f), C oracle gcc -O2, 1 thread, %.1f s"
                     % (len(pick), ",".join(str(int(i)) for i in pick), nrhs / len(pick), st["n_rhs"].mean(), el)}
    # the host PES path the surrogate replaces: the oracle's site-site
    # potential on one core (C loop over an 8^5 x 8 grid), scaled to the
    # 16^5 x 128 grid
    try:
        from mqct_amd import workloads as W
        w3 = W.two_asym_tops()
        ws = W.two_asym_tops(jmax1=0, jmax2=0, n_ang=8, n_r=8)
        t0 = time.perf_counter()
        O.oracle_for_workload(ws)
        per = (time.perf_counter() - t0) / (ws.grid.G * len(ws.R))
        out["host_pes_s_per_point"] = per
        out["host_pes_grid_s_est"] = per * w3.grid.G * len(w3.R)
        log("cpu oracle PES: %.3f us per point -> %.0f s for the two-tops grid" % (
            per * 1e6, out["host_pes_grid_s_est"]))
    except Exception as e:  # noqa: BLE001
        log("host PES timing skipped:", e)
    return out


def baseline_sample(n_rhs, n_sample):
    """Indices of the n_sample trajectories whose right-hand-side counts are
    closest to the batch mean (ties by index)."""
    n_rhs = np.asarray(n_rhs, dtype=np.float64)
    d = np.abs(n_rhs - n_rhs.mean())
    return [int(i) for i in np.argsort(d, kind="stable")[:min(n_sample, len(n_rhs))]]


if __name__ == "__main__":
    sys.exit(main())
