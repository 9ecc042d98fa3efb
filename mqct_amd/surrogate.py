"""The learned potential for the two-asymmetric-tops system (SYS_TYPE 0):
an MLP 12 -> 256 x 4 -> 1 (SiLU, BF16 weights, FP32 accumulation) trained
and evaluated by the library's HIP kernels.  It stands where the reference
calls the user PES on the angular grid (pes_sys_type.f:88-97, through
INI_POT_STORAGE); the plain potential stays the fallback below r_min and
the reference against which the network's error is measured.

    sur = Surrogate(dev, seed=1)
    geo = sample_geometries(4_000_000, 4.0, 25.0, seed=...)
    V = pes_points(dev, wl, geo)                 # the method's own PES
    sur.fit(geo, V, steps=..., batch=...)
    mb = MatrixBuilder(dev, wl)                  # plain V grid (fallback)
    sur.fill_vgrid(wl.R)                         # V grid from the network
    M, D, S = mb.build("bf16x3")

The network predicts y with V = scale*sinh(y) + shift (cm^-1): relative
accuracy on the repulsive wall, absolute accuracy at long range.
"""
import ctypes

import numpy as np

from . import _lib as L


def sample_geometries(n, r_lo, r_hi, seed, r_power=1.0):
    """n random geometries (R, beta1, gamma1, alpha2, beta2, gamma2): R in
    [r_lo, r_hi] with density ~ R^(-r_power) (more points where the
    potential varies fast), cos(beta) uniform, the other angles uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(0.0, 1.0, n)
    if r_power == 1.0:
        R = r_lo * (r_hi / r_lo) ** u
    else:
        a = 1.0 - r_power
        R = (r_lo ** a + u * (r_hi ** a - r_lo ** a)) ** (1.0 / a)
    b1 = np.arccos(rng.uniform(-1.0, 1.0, n))
    b2 = np.arccos(rng.uniform(-1.0, 1.0, n))
    g1, a2, g2 = (rng.uniform(0.0, 2 * np.pi, n) for _ in range(3))
    return np.ascontiguousarray(np.column_stack([R, b1, g1, a2, b2, g2]))


def grid_geometries(wl):
    """The (R, angle) points of a SYS_TYPE 0 workload's V grid in the
    POT_STORE order of INI_POT_STORAGE (R outermost, then alpha2, beta1,
    beta2, gamma1, gamma2)."""
    g = wl.grid
    b1 = g.xbm + g.xbl * np.asarray(g.xb)
    b2 = g.xbbm + g.xbbl * np.asarray(g.xbb)
    g1, a2, g2 = np.asarray(g.xg), np.asarray(g.xaa), np.asarray(g.xgg)
    R, A2, B1, B2, G1, G2 = np.meshgrid(wl.R, a2, b1, b2, g1, g2, indexing="ij")
    return np.ascontiguousarray(np.column_stack([R.ravel(), B1.ravel(), G1.ravel(), A2.ravel(), B2.ravel(),
                                                 G2.ravel()]))


def pes_points(dev, wl, geo):
    """The workload's synthetic potential at arbitrary geometries, evaluated
    by the library's PES kernel (the same code that fills the plain V grid)."""
    geo = np.ascontiguousarray(geo, np.float64)
    par = L.as_f64(wl.pes_par)
    V = np.zeros(len(geo))
    CH = 1 << 24
    for s in range(0, len(geo), CH):
        g = np.ascontiguousarray(geo[s:s + CH])
        L.check(dev.lib.mqct_hip_pes_points(dev.h, wl.pes_kind, len(par), L.dptr(par), len(g), L.dptr(g),
                                            L.dptr(V[s:s + CH])), "pes_points")
    return V


class Surrogate:
    def __init__(self, dev, seed=1, scale=100.0, shift=0.0, r_min=4.0):
        self.dev, self.lib, self.h = dev, dev.lib, dev.h
        L.check(self.lib.mqct_hip_mlp_init(self.h, seed), "mlp_init")
        self.n_params = self.lib.mqct_hip_mlp_n_params()
        self.set_scaling(scale, shift, r_min)
        self.n_data = 0

    def set_scaling(self, scale, shift, r_min):
        self.scale, self.shift, self.r_min = scale, shift, r_min
        L.check(self.lib.mqct_hip_mlp_set_scaling(self.h, scale, shift, r_min), "mlp_set_scaling")

    def target(self, V):
        return np.arcsinh((np.asarray(V) - self.shift) / self.scale).astype(np.float32)

    def potential(self, y):
        return self.scale * np.sinh(np.asarray(y, np.float64)) + self.shift

    # -- parameters -----------------------------------------------------------
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        L.check(self.lib.mqct_hip_mlp_get_params(self.h, p.ctypes.data_as(L.c_flt_p)), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        L.check(self.lib.mqct_hip_mlp_set_params(self.h, p.ctypes.data_as(L.c_flt_p)), "mlp_set_params")

    # -- training -------------------------------------------------------------
    def set_data(self, geo, V=None, y=None):
        geo = np.ascontiguousarray(geo, np.float64)
        y = self.target(V) if y is None else np.ascontiguousarray(y, np.float32)
        L.check(self.lib.mqct_hip_mlp_set_data(self.h, len(geo), L.dptr(geo), y.ctypes.data_as(L.c_flt_p)),
                "mlp_set_data")
        self.n_data = len(geo)

    def grad(self, start, n):
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        L.check(self.lib.mqct_hip_mlp_grad(self.h, start, n, g.ctypes.data_as(L.c_flt_p), ctypes.byref(loss)),
                "mlp_grad")
        return g, loss.value

    def step(self, start, n, lr):
        loss = ctypes.c_float()
        L.check(self.lib.mqct_hip_mlp_train_step(self.h, start, n, lr, ctypes.byref(loss)), "mlp_train_step")
        return loss.value

    def fit(self, steps, batch, lr0=1e-3, lr1=1e-5, log_every=0, log=None):
        """Adam over consecutive batches of the (pre-shuffled) training set
        with a cosine learning-rate decay from lr0 to lr1."""
        nb = max(1, self.n_data // batch)
        losses = []
        for s in range(steps):
            lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + np.cos(np.pi * s / max(1, steps - 1)))
            b = s % nb
            losses.append(self.step(b * batch, min(batch, self.n_data - b * batch), lr))
            if log and log_every and (s % log_every == 0 or s == steps - 1):
                log("step %d lr %.2e loss %.3e" % (s, lr, losses[-1]))
        return np.array(losses)

    # -- inference ------------------------------------------------------------
    def predict_y(self, geo):
        geo = np.ascontiguousarray(geo, np.float64)
        y = np.zeros(len(geo), np.float32)
        CH = 1 << 20
        for s in range(0, len(geo), CH):
            g = np.ascontiguousarray(geo[s:s + CH])
            L.check(self.lib.mqct_hip_mlp_predict(self.h, len(g), L.dptr(g),
                                                  y[s:s + CH].ctypes.data_as(L.c_flt_p)), "mlp_predict")
        return y

    def predict(self, geo):
        return self.potential(self.predict_y(geo))

    def fill_vgrid(self, R):
        """Replace the V grid of the current SYS_TYPE 0 grid by the network's
        (points below r_min keep the plain potential already there)."""
        R = L.as_f64(R)
        L.check(self.lib.mqct_hip_surrogate_vgrid(self.h, len(R), L.dptr(R)), "surrogate_vgrid")


def train_surrogate(dev, wl, n_samples=4_000_000, steps=20000, batch=4096, seed=20240506, r_lo=None,
                    r_hi=None, lr0=2e-3, lr1=2e-5, scale=200.0, log=None):
    """Trains a surrogate of the workload's potential on n_samples seeded
    random geometries that the library's own PES kernel evaluates (R with
    density ~ 1/R over [r_lo, r_hi], by default the workload's R range; grid
    points below r_lo keep the plain potential).  Returns the network and a
    dict of timings and fit errors on a held-out set.  The output scale of
    200 cm^-1 keeps the attractive well and the low wall in the linear
    range of the sinh transform (on the two-tops system it gave cross
    sections 2.6x closer to the plain-PES ones than 20 cm^-1)."""
    import time
    r_lo = float(wl.R[0]) if r_lo is None else r_lo
    r_hi = float(wl.R[-1]) if r_hi is None else r_hi
    t0 = time.perf_counter()
    geo = sample_geometries(n_samples, r_lo, r_hi, seed)
    V = pes_points(dev, wl, geo)
    t_data = time.perf_counter() - t0
    sur = Surrogate(dev, seed=seed, scale=scale, shift=0.0, r_min=r_lo)
    sur.set_data(geo, V)
    t0 = time.perf_counter()
    losses = sur.fit(steps, batch, lr0=lr0, lr1=lr1, log_every=max(1, steps // 10), log=log)
    t_fit = time.perf_counter() - t0
    test = sample_geometries(200_000, r_lo, r_hi, seed + 1)
    Vt = pes_points(dev, wl, test)
    Vp = sur.predict(test)
    err = np.abs(Vp - Vt)
    big = np.abs(Vt) > 100.0
    stats = {"samples": n_samples, "steps": steps, "batch": batch, "data_s": t_data, "fit_s": t_fit,
             "final_loss": float(np.mean(losses[-100:])),
             "test_abs_err_small_V_cm": float(err[~big].max()) if (~big).any() else None,
             "test_rel_err_large_V": float((err[big] / np.abs(Vt[big])).max()) if big.any() else None,
             "test_rms_abs_err_cm": float(np.sqrt(np.mean(err ** 2)))}
    return sur, stats

