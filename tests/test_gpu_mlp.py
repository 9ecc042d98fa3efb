"""GPU checks of the PES surrogate network (HIP forward, backward and Adam
kernels) against the FP64 numpy oracle of the same network.

Bars: the oracle with BF16 rounding at the GPU's storage points
(quant=True) leaves only FP32 accumulation order and rare BF16 rounding
flips, so the MSE gradient agrees per parameter block to 2% of the block's
norm and the forward output to 1e-2 of the output scale; the loss within
1e-3 relative.  Training must reduce the loss on the two-tops potential.
"""
import numpy as np
import pytest

from mqct_amd import workloads as W
from mqct_amd.matrix import MatrixBuilder
from mqct_amd.surrogate import Surrogate, grid_geometries, pes_points, sample_geometries
import oracle.mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data(dev):
    wl = W.two_asym_tops()
    geo = sample_geometries(4096, 4.0, 25.0, seed=21)
    V = pes_points(dev, wl, geo)
    return wl, geo, V


def test_pes_points_match_oracle_potential(dev, data):
    """The GPU PES kernel that makes the training targets against the
    oracle's potential (the same function that fills the plain V grid)."""
    import ctypes
    import oracle.oracle as O
    wl, geo, V = data
    par = np.ascontiguousarray(wl.pes_par, np.float64)
    f = O.lib().or_pes_sites
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.POINTER(ctypes.c_double)] + [ctypes.c_double] * 6
    ref = np.array([f(par.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), *g) for g in geo[:512]])
    # FP64 on both sides; the GPU contracts multiply-adds, the oracle does not
    assert np.max(np.abs(V[:512] - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-11


def test_gradient_matches_oracle(dev, data):
    wl, geo, V = data
    sur = Surrogate(dev, seed=11)
    sur.set_data(geo, V)
    n = 1024
    g, loss = sur.grad(0, n)
    p = sur.params().astype(np.float64)
    X = MO.features(geo[:n])
    t = sur.target(V[:n]).astype(np.float64)
    Lq, gq = MO.loss_grad(p, X, t, quant=True)
    assert abs(loss - Lq) <= 1e-3 * Lq
    for name, sl in MO.block_slices().items():
        ref = gq[sl]
        err = np.linalg.norm(g[sl] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert err <= 2e-2, f"{name}: relative gradient error {err:.3e}"
    # and the exact FP64 gradient is within BF16 storage error of both
    L0, g0 = MO.loss_grad(p, X, t)
    assert np.linalg.norm(g - g0) <= 5e-2 * np.linalg.norm(g0)


def test_forward_matches_oracle(dev, data):
    wl, geo, V = data
    sur = Surrogate(dev, seed=12)
    y = sur.predict_y(geo[:2048])
    yq, _ = MO.forward(sur.params().astype(np.float64), MO.features(geo[:2048]), quant=True)
    assert np.max(np.abs(y - yq)) <= 1e-2 * max(1.0, np.abs(yq).max())


def test_training_reduces_loss(dev, data):
    wl, geo, V = data
    sur = Surrogate(dev, seed=13)
    sur.set_data(geo, V)
    _, l0 = sur.grad(0, len(geo))
    losses = sur.fit(steps=400, batch=512, lr0=2e-3, lr1=1e-4)
    _, l1 = sur.grad(0, len(geo))
    assert np.all(np.isfinite(losses))
    assert l1 < 0.05 * l0, (l0, l1)


def test_surrogate_vgrid_is_the_network_on_the_grid(dev):
    """fill_vgrid writes scale*sinh(y)+shift of the network at every grid
    point in POT_STORE order: a build from it equals a build from the host
    V grid made by predict() on grid_geometries()."""
    wl = W.two_asym_tops(jmax1=1, jmax2=1, n_ang_each=(3, 4, 3, 4, 3), n_r=6)
    sur = Surrogate(dev, seed=14, scale=100.0, shift=0.0, r_min=0.0)
    Vh = sur.predict(grid_geometries(wl))
    mb = MatrixBuilder(dev, wl, vgrid=Vh)
    M1, D1, S1 = mb.build("fp64")
    mb2 = MatrixBuilder(dev, wl)
    sur.fill_vgrid(wl.R)
    M2, D2, S2 = mb2.build("fp64")
    assert np.array_equal(S1, S2)
    scale = np.abs(M1).max()
    assert np.max(np.abs(M1 - M2)) <= 1e-12 * scale


def test_surrogate_gain_over_host_pes(dev, data):
    """What the network buys: the potential on two-tops geometries from the
    trained network on the GPU, against the host's PES routine (the
    oracle's C site-site potential filling an 8^5 x 8 grid on one core, as
    the reference fills its grid point by point).  The network must be at
    least 10x faster per point (measured: about 80x); its error on random
    points is printed beside the timing."""
    import time
    import oracle.oracle as O
    wl, geo, V = data
    sur = Surrogate(dev, seed=15, scale=20.0)
    sur.set_data(geo, V)
    sur.fit(steps=400, batch=512, lr0=2e-3, lr1=1e-4)
    pts = sample_geometries(262144, 5.0, 25.0, seed=23)
    sur.predict(pts[:4096])  # first call: kernels loaded
    t0 = time.perf_counter()
    Vn = sur.predict(pts)
    t_net = (time.perf_counter() - t0) / len(pts)
    ws = W.two_asym_tops(jmax1=0, jmax2=0, n_ang=8, n_r=8)
    t0 = time.perf_counter()
    O.oracle_for_workload(ws)
    t_host = (time.perf_counter() - t0) / (ws.grid.G * len(ws.R))
    Vp = pes_points(dev, wl, pts[:20000])
    err = np.abs(Vn[:20000] - Vp)
    print(f"network {t_net * 1e9:.1f} ns/point, host PES {t_host * 1e9:.1f} ns/point "
          f"(x{t_host / t_net:.0f}); |V_net - V| rms {np.sqrt(np.mean(err ** 2)):.1f} cm^-1")
    assert np.all(np.isfinite(Vn))
    assert t_host >= 10.0 * t_net
