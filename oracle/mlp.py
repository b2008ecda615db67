"""Float64 oracle of the per-sub-aperture MLP centroider (test infrastructure
only).  No counterpart in the reference: the network is this build's learned
centroider.

Network 64 -> 128 -> 128 -> 2, GELU (tanh form, Hendrycks & Gimpel 2016):
    gelu(z)  = 0.5 z (1 + tanh(c (z + 0.044715 z^3))),  c = sqrt(2 / pi)
    z1 = W1 x + b1, h1 = gelu(z1); z2 = W2 h1 + b2, h2 = gelu(z2); y = W3 h2 + b3
Loss: mean over samples and outputs of (y - t)^2.

Flat parameter vector (the device layout): W1 (128 x 64, row-major),
b1 (128), W2 (128 x 128), b2 (128), W3 (2 x 128), b3 (2).
"""
import numpy as np

NIN, NH, NOUT = 64, 128, 2
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NOUT, NH)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
C_GELU = np.sqrt(2.0 / np.pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = C_GELU * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def forward(p, x):
    """x (n, 64) -> y (n, 2) and the cached activations."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = gelu(z1)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = gelu(z2)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (x, z1, h1, z2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_grad(p, x, t):
    """Mean-squared-error loss and its gradient w.r.t. the flat parameters."""
    w = unpack(p)
    y, (x, z1, h1, z2, h2) = forward(p, x)
    t = np.asarray(t, dtype=np.float64)
    n = x.shape[0]
    dy = 2.0 * (y - t) / (n * NOUT)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dz2 = (dy @ w["W3"]) * gelu_grad(z2)
    g["W2"], g["b2"] = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ w["W2"]) * gelu_grad(z1)
    g["W1"], g["b1"] = dz1.T @ x, dz1.sum(0)
    return float(np.mean((y - t) ** 2)), pack(g)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) with bias correction; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
