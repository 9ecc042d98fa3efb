"""FP64 numpy restatement of the PES surrogate network (12 -> 256 x 4 -> 1,
SiLU) and of its mean-squared-error gradient.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the HIP training and
inference kernels against; the product never imports it.

The network has no counterpart in the reference (the reference evaluates
its PES routines directly, pot_comp.f / pes_sys_type.f); what this file
pins is the calculus: tests/test_mlp_oracle.py checks the analytic gradient
against central finite differences in FP64.  With quant=True the forward
and backward passes round to BF16 wherever the GPU kernels store BF16
(features, weights of layers 1-4, hidden activations, back-propagated
pre-activation gradients), so the remaining GPU difference is FP32
accumulation order.

Parameter layout (flat, float): W1[256][12] b1[256] W2[256][256] b2 W3 b3
W4 b4 W5[256] b5[1].
"""
import numpy as np

NF, HW, NH = 12, 256, 4
SIZES = [("W1", (HW, NF)), ("b1", (HW,)), ("W2", (HW, HW)), ("b2", (HW,)), ("W3", (HW, HW)), ("b3", (HW,)),
         ("W4", (HW, HW)), ("b4", (HW,)), ("W5", (HW,)), ("b5", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) and return as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.ravel(d[name]) for name, _ in SIZES])


def features(geo):
    """geo [n][6] = (R, beta1, gamma1, alpha2, beta2, gamma2) ->
    [n][12] = (4/R, exp(-(R-4)/4), cos, sin of the five angles)."""
    geo = np.asarray(geo, np.float64)
    R = geo[:, 0]
    f = [4.0 / R, np.exp(-(R - 4.0) / 4.0)]
    for a in range(5):
        f += [np.cos(geo[:, 1 + a]), np.sin(geo[:, 1 + a])]
    return np.stack(f, 1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, quant=False):
    """Returns y [n] and the cache (Z_l, H_l)."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    H = q(X)
    Hs, Zs = [H], []
    for l in range(NH):
        W, b = q(P["W%d" % (l + 1)]), P["b%d" % (l + 1)]
        Z = H @ W.T + b
        H = q(silu(Z))
        Zs.append(Z)
        Hs.append(H)
    y = H @ P["W5"] + P["b5"][0]
    return y, (Zs, Hs)


def loss_grad(p, X, t, quant=False):
    """L = mean((y - t)^2) and dL/dp in the flat layout."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    y, (Zs, Hs) = forward(p, X, quant)
    n = len(t)
    e = y - np.asarray(t, np.float64)
    L = float(np.mean(e * e))
    dy = 2.0 * e / n
    G = {"W5": Hs[NH].T @ dy, "b5": np.array([dy.sum()])}
    dZ = q(dy[:, None] * P["W5"][None, :] * dsilu(Zs[NH - 1]))
    for l in range(NH - 1, -1, -1):
        G["W%d" % (l + 1)] = dZ.T @ Hs[l]
        G["b%d" % (l + 1)] = dZ.sum(0)
        if l > 0:
            W = q(P["W%d" % (l + 1)])
            dZ = q((dZ @ W) * dsilu(Zs[l - 1]))
    return L, pack(G)


def block_slices():
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = slice(o, o + n)
        o += n
    return out
