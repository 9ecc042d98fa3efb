"""The surrogate network's FP64 oracle: gradient against central finite
differences (the calculus the GPU backward pass is checked against), and
the BF16 rounding helper against exactly representable cases."""
import numpy as np

import oracle.mlp_oracle as MO


def _setup(n=16, seed=3):
    rng = np.random.default_rng(seed)
    p = np.zeros(MO.NPARAM)
    S = MO.block_slices()
    fan = {"W1": MO.NF, "W2": MO.HW, "W3": MO.HW, "W4": MO.HW, "W5": MO.HW}
    for k, sl in S.items():
        sd = np.sqrt(2.0 / fan[k]) if k in fan else 0.1
        p[sl] = rng.normal(0, sd, sl.stop - sl.start)
    geo = np.column_stack([rng.uniform(3.5, 12.0, n), rng.uniform(0, np.pi, n),
                           *[rng.uniform(0, 2 * np.pi, n) for _ in range(4)]])
    X = MO.features(geo)
    t = rng.normal(0, 1, n)
    return p, X, t


def test_gradient_matches_finite_differences():
    p, X, t = _setup()
    L, g = MO.loss_grad(p, X, t)
    rng = np.random.default_rng(5)
    S = MO.block_slices()
    # a few parameters of every block, including both biases at the ends
    idx = [S["b5"].start, S["b1"].start + 7]
    for k, sl in S.items():
        idx += list(rng.integers(sl.start, sl.stop, 4))
    for i in idx:
        h = 1e-5 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (MO.loss_grad(pp, X, t)[0] - MO.loss_grad(pm, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()), (i, fd, g[i])


def test_quantised_gradient_is_close_to_exact():
    """BF16 storage perturbs the gradient by ~2^-8 relative, not more."""
    p, X, t = _setup(n=64)
    _, g = MO.loss_grad(p, X, t)
    _, gq = MO.loss_grad(p, X, t, quant=True)
    assert np.linalg.norm(gq - g) <= 3e-2 * np.linalg.norm(g)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0])
    # 8 significant bits: 1+2^-7 is representable; 1+2^-8 is a tie -> even
    # (1.0); 1+3*2^-8 is a tie between 1+2^-7 (odd) and 1+2^-6 (even)
    assert np.array_equal(MO.bf16(x), [1.0, 1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -6, -2.5, 0.0])
