"""Pins the FP64 numpy reference of the closure network (oracle/closure_ref.py)
by central finite differences and by a direct-sum convolution."""
import numpy as np

import closure_ref as cr


def small_net(rng, chans=(4, 6, 5, 3)):
    return [(0.3 * rng.standard_normal((9, a, b)), 0.1 * rng.standard_normal(b))
            for a, b in zip(chans[:-1], chans[1:])]


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 7, 3))
    w = rng.standard_normal((9, 3, 4))
    b = rng.standard_normal(4)
    z = cr.conv_fwd(x, w, b)
    n, h, wd, _ = x.shape
    ref = np.zeros((n, h, wd, 4))
    for i in range(n):
        for y in range(h):
            for xx in range(wd):
                acc = b.copy()
                for t, (dy, dx) in enumerate(cr.TAPS):
                    yy, x2 = y + dy, xx + dx
                    if 0 <= yy < h and 0 <= x2 < wd:
                        acc += x[i, yy, x2] @ w[t]
                ref[i, y, xx] = acc
    assert np.allclose(z, ref, rtol=1e-13, atol=1e-13)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = small_net(rng)
    x = rng.standard_normal((2, 6, 5, 4))
    t = rng.standard_normal((2, 6, 5, 3))
    wch = np.array([1.0, 0.5, 2.0])
    loss, grads = cr.loss_and_grads(params, x, t, wch)
    h = 1e-6
    for k, (w, b) in enumerate(params):
        for arr, g in ((w, grads[k][0]), (b, grads[k][1])):
            flat = arr.reshape(-1)
            gflat = g.reshape(-1)
            for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
                old = flat[idx]
                flat[idx] = old + h
                lp, _ = cr.loss_and_grads(params, x, t, wch)
                flat[idx] = old - h
                lm, _ = cr.loss_and_grads(params, x, t, wch)
                flat[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - gflat[idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, gflat[idx])


def test_input_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    w = rng.standard_normal((9, 3, 2))
    x = rng.standard_normal((1, 4, 4, 3))
    dz = rng.standard_normal((1, 4, 4, 2))
    dx, _, _ = cr.conv_bwd(x, w, dz)
    h = 1e-6
    for idx in np.ndindex(x.shape):
        xp = x.copy(); xp[idx] += h
        xm = x.copy(); xm[idx] -= h
        fd = (np.sum(dz * cr.conv_fwd(xp, w, 0)) - np.sum(dz * cr.conv_fwd(xm, w, 0))) / (2 * h)
        assert abs(fd - dx[idx]) <= 1e-6 * max(1.0, abs(fd))
