"""FP64 numpy reference of the closure network -- test infrastructure only.

The network the library trains on the GPU with its own BF16-MFMA kernels
(nemolite2d-cuda_amd/csrc/nemo2d_closure.hip): 'same'-padded 3x3
convolutions on NHWC tensors, exact (erf) GELU between layers, weighted MSE
loss.  Plain loops over the 9 taps; used by the tests to check the GPU
forward pass and gradients, and itself pinned by finite differences
(tests/test_closure_ref.py).
"""
import numpy as np
from scipy.special import erf

TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]  # tap = 3*(dy+1) + (dx+1)


def shift(x, dy, dx):
    """out[n, y, x, c] = x[n, y+dy, x+dx, c] (zero outside)."""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys = slice(max(0, -dy), min(h, h - dy))
    xs = slice(max(0, -dx), min(w, w - dx))
    yd = slice(max(0, dy), min(h, h + dy))
    xd = slice(max(0, dx), min(w, w + dx))
    out[:, ys, xs, :] = x[:, yd, xd, :]
    return out


def conv_fwd(x, w, b):
    """x [n,h,w,cin], w [9,cin,cout], b [cout] -> z [n,h,w,cout]."""
    z = np.zeros(x.shape[:3] + (w.shape[2],))
    for t, (dy, dx) in enumerate(TAPS):
        z += shift(x, dy, dx) @ w[t]
    return z + b


def conv_bwd(x, w, dz):
    """Gradients of sum(dz * conv_fwd(x, w, b)) w.r.t. x, w, b."""
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, (dy, ddx) in enumerate(TAPS):
        dx += shift(dz @ w[t].T, -dy, -ddx)
        xs = shift(x, dy, ddx)
        dw[t] = np.einsum("nhwc,nhwd->cd", xs, dz)
    db = dz.sum(axis=(0, 1, 2))
    return dx, dw, db


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, x):
    """params: list of (w, b); GELU after every layer but the last.
    Returns (output, cache)."""
    cache = []
    a = x
    for k, (w, b) in enumerate(params):
        z = conv_fwd(a, w, b)
        cache.append((a, z))
        a = gelu(z) if k < len(params) - 1 else z
    return a, cache


def loss_and_grads(params, x, target, wch):
    """Weighted MSE: mean over pixels and channels of wch[c]*(y-t)^2."""
    y, cache = forward(params, x)
    d = y - target
    count = d.size
    loss = float(np.sum(wch * d * d) / count)
    g = 2.0 * wch * d / count
    grads = [None] * len(params)
    for k in range(len(params) - 1, -1, -1):
        a, z = cache[k]
        if k < len(params) - 1:
            g = g * gelu_grad(z)
        dx, dw, db = conv_bwd(a, params[k][0], g)
        grads[k] = (dw, db)
        g = dx
    return loss, grads
