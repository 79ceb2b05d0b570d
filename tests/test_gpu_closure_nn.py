"""The closure network's GPU kernels (BF16 MFMA implicit-GEMM convolutions,
GELU, MSE, Adam) against the FP64 numpy reference oracle/closure_ref.py."""
import numpy as np
import pytest

import closure_ref as cr

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round FP64 values to BF16 (nearest even), back to FP64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def setup(seed, n=2, h=11, w=13):
    from nemo2d.closure import ConvNet, LAYERS
    rng = np.random.default_rng(seed)
    net = ConvNet(seed=seed)
    ws = [0.5 * rng.standard_normal((9, ci, co)) / np.sqrt(9 * ci) for ci, co, _ in LAYERS]
    bs = [0.1 * rng.standard_normal(co) for _, co, _ in LAYERS]
    net.set_weights(ws, bs)
    x = rng.standard_normal((n, h, w, 4))
    t = rng.standard_normal((n, h, w, 3)) * 0.1
    xp = np.zeros((n * h * w, 8), dtype=np.float32)
    xp[:, :4] = x.reshape(-1, 4)
    tp = np.zeros((n * h * w, 8), dtype=np.float32)
    tp[:, :3] = t.reshape(-1, 3)
    xg = net.to_bf16(xp)
    tg = net.upload(tp)
    # the reference sees the BF16-rounded operands the GPU sees
    params = [(bf16(w_), b_.astype(np.float32).astype(np.float64)) for w_, b_ in zip(ws, bs)]
    return net, params, bf16(x), t, xg, tg, (n, h, w)


def relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_forward_matches_reference():
    net, params, x, t, xg, tg, (n, h, w) = setup(0)
    y = net.forward(xg, n, h, w).download()[:, :3].reshape(n, h, w, 3)
    ref, _ = cr.forward(params, x)
    assert relerr(y, ref) < 2e-2, relerr(y, ref)


def test_inference_forward_equals_training_forward():
    """forward(train=False) skips the hidden pre-activation stores only: the
    output is bitwise the training forward's."""
    net, params, x, t, xg, tg, (n, h, w) = setup(4, n=3, h=29, w=45)
    y_train = net.forward(xg, n, h, w).download()
    y_inf = net.forward(xg, n, h, w, train=False).download()
    assert np.array_equal(y_train, y_inf)


def test_first_layer_exact_accumulation():
    """One convolution with BF16 operands: FP32 accumulation of exactly
    representable products; agrees with FP64 to FP32 rounding."""
    from nemo2d import closure as C
    rng = np.random.default_rng(3)
    n, h, w, cin, cout = 1, 7, 9, 4, 32
    net = C.ConvNet(seed=3, fp32_preact=True)
    wgt = bf16(rng.standard_normal((9, cin, cout)))
    bias = rng.standard_normal(cout).astype(np.float32).astype(np.float64)
    net.set_weights([wgt] + [np.zeros((9, a, b)) for a, b, _ in C.LAYERS[1:]],
                    [bias] + [np.zeros(b) for _, b, _ in C.LAYERS[1:]])
    x = bf16(rng.standard_normal((n, h, w, cin)))
    xp = np.zeros((n * h * w, 8), dtype=np.float32)
    xp[:, :4] = x.reshape(-1, 4)
    net.forward(net.to_bf16(xp), n, h, w)
    z0 = net.z[0].download()[:, :cout].reshape(n, h, w, cout)
    ref = cr.conv_fwd(x, wgt, bias)
    assert np.allclose(z0, ref, rtol=2e-6, atol=2e-5)


def test_gradients_match_reference():
    net, params, x, t, xg, tg, (n, h, w) = setup(1)
    wch = net.upload(np.array([1.0, 0.5, 2.0]))
    net.forward(xg, n, h, w)
    loss = float(net.backward(tg, wch))
    gw, gb = net.grads()
    ref_loss, ref_grads = cr.loss_and_grads(params, x, t, np.array([1.0, 0.5, 2.0]))
    assert abs(loss - ref_loss) <= 3e-2 * ref_loss
    for k in range(4):
        assert relerr(gw[k], ref_grads[k][0]) < 5e-2, (k, relerr(gw[k], ref_grads[k][0]))
        assert relerr(gb[k], ref_grads[k][1]) < 5e-2, (k, relerr(gb[k], ref_grads[k][1]))


def test_gradients_match_reference_over_several_pixel_chunks():
    """3 images of 41 x 43 pixels: a batch that is not a multiple of the
    weight gradient's 4-image groups, tiles cut by the image edges, the
    weight gradient summed over many blocks and the loss over many blocks."""
    net, params, x, t, xg, tg, (n, h, w) = setup(4, n=3, h=41, w=43)
    wch = net.upload(np.array([1.0, 1.0, 1.0]))
    net.forward(xg, n, h, w)
    loss = float(net.backward(tg, wch))
    gw, gb = net.grads()
    ref_loss, ref_grads = cr.loss_and_grads(params, x, t, np.ones(3))
    assert abs(loss - ref_loss) <= 3e-2 * ref_loss
    for k in range(4):
        assert relerr(gw[k], ref_grads[k][0]) < 5e-2, (k, relerr(gw[k], ref_grads[k][0]))
        assert relerr(gb[k], ref_grads[k][1]) < 5e-2, (k, relerr(gb[k], ref_grads[k][1]))


def test_forward_and_gradients_with_several_tiles_per_block():
    """6 images of 130 x 260 pixels: 918 tiles of 8 x 32, more than the
    convolution kernels' resident blocks (512 for 32 input channels, 768
    for the first layer's 16), so each block walks several tiles through
    its two LDS buffers; tiles cut by both image edges."""
    net, params, x, t, xg, tg, (n, h, w) = setup(7, n=6, h=130, w=260)
    y = net.forward(xg, n, h, w).download()[:, :3].reshape(n, h, w, 3)
    ref, _ = cr.forward(params, x)
    assert relerr(y, ref) < 2e-2, relerr(y, ref)
    wch = net.upload(np.ones(3))
    loss = float(net.backward(tg, wch))
    gw, gb = net.grads()
    ref_loss, ref_grads = cr.loss_and_grads(params, x, t, np.ones(3))
    assert abs(loss - ref_loss) <= 3e-2 * ref_loss
    for k in range(4):
        assert relerr(gw[k], ref_grads[k][0]) < 5e-2, (k, relerr(gw[k], ref_grads[k][0]))
        assert relerr(gb[k], ref_grads[k][1]) < 5e-2, (k, relerr(gb[k], ref_grads[k][1]))


@pytest.mark.parametrize("fp32_preact", [False, True])
def test_gradients_with_either_preactivation_storage(fp32_preact):
    """The GELU backward reads the layer below's pre-activation stored in
    BF16 (the default) or FP32: both give the reference's gradients."""
    from nemo2d.closure import ConvNet
    net, params, x, t, xg, tg, (n, h, w) = setup(6, n=5, h=37, w=70)
    ws, bs = net.weights()
    net2 = ConvNet(seed=6, fp32_preact=fp32_preact)
    net2.set_weights(ws, bs)
    wch = net2.upload(np.ones(3))
    net2.forward(xg, n, h, w)
    loss = float(net2.backward(tg, wch))
    gw, gb = net2.grads()
    ref_loss, ref_grads = cr.loss_and_grads(params, x, t, np.ones(3))
    assert abs(loss - ref_loss) <= 3e-2 * ref_loss
    for k in range(4):
        assert relerr(gw[k], ref_grads[k][0]) < 5e-2, (k, relerr(gw[k], ref_grads[k][0]))
        assert relerr(gb[k], ref_grads[k][1]) < 5e-2, (k, relerr(gb[k], ref_grads[k][1]))


def bf16_bits_to_f64(a):
    return (np.asarray(a).astype(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_gelu_and_its_derivative_against_erf():
    """The kernels' GELU (forward epilogue) and GELU' (nemo2d_nn_gelu_bwd)
    over pre-activations in [-8, 8] against z/2 (1 + erf(z/sqrt2)) and
    Phi(z) + z phi(z) in FP64: within BF16 rounding (2^-8 relative, one
    rounding and the FP32 arithmetic) plus 1e-6 absolute (the erf
    approximation's 1.5e-7 times |z|/2)."""
    from math import erf, exp, pi, sqrt
    from nemo2d import closure as C
    rng = np.random.default_rng(11)
    n, h, w, cin, cout = 1, 16, 40, 4, 32
    net = C.ConvNet(seed=11, fp32_preact=True)
    wgt = np.zeros((9, cin, cout))
    wgt[4, 0, :] = 1.0  # z[p][co] = x[p][0] + b[co]
    bias = np.arange(cout) * 0.375 - 6.0
    net.set_weights([wgt] + [np.zeros((9, a, b)) for a, b, _ in C.LAYERS[1:]],
                    [bias] + [np.zeros(b) for _, b, _ in C.LAYERS[1:]])
    x0 = bf16(rng.uniform(-2.0, 2.0, (n * h * w)))
    xp = np.zeros((n * h * w, 8), dtype=np.float32)
    xp[:, 0] = x0
    net.forward(net.to_bf16(xp), n, h, w)
    z = net.z[0].download()[:, :cout].astype(np.float64)
    assert np.array_equal(z, x0[:, None] + bias[None, :])  # exact sums
    act = bf16_bits_to_f64(net.act[0].download()[:, :cout])
    ref = np.vectorize(lambda v: 0.5 * v * (1.0 + erf(v / sqrt(2.0))))(z)
    err = np.abs(act - ref) - (2.0 ** -8 * np.abs(ref) + 1e-6)
    assert err.max() <= 0, (err.max(), z.ravel()[np.argmax(err)])
    zf = C.DevArray.from_numpy(z.astype(np.float32).ravel())
    one = C.DevArray.from_numpy(np.ones(z.size, dtype=np.float32))
    dz = C.DevArray((z.size,), np.int16)
    C._check(net.L.nemo2d_nn_gelu_bwd(one.fptr(), zf.fptr(), z.size, dz.ptr, None))
    d = bf16_bits_to_f64(dz.download()).reshape(z.shape)
    dref = np.vectorize(lambda v: 0.5 * (1.0 + erf(v / sqrt(2.0))) + v * exp(-0.5 * v * v) / sqrt(2 * pi))(z)
    derr = np.abs(d - dref) - (2.0 ** -8 * np.abs(dref) + 1e-6)
    assert derr.max() <= 0, (derr.max(), z.ravel()[np.argmax(derr)])
    for a in (zf, one, dz):
        a.close()


def test_device_arrays_round_trip():
    from nemo2d.closure import DevArray
    a = np.arange(1000, dtype=np.float32) * 0.5
    d = DevArray.from_numpy(a)
    assert np.array_equal(d.download(), a)
    d.zero()
    assert not d.download().any()
    b = DevArray((7, 3), np.int16)
    assert b.download().shape == (7, 3) and not b.download().any()
    d.close()
    b.close()


def test_training_reduces_loss():
    """Adam on a fixed batch: the loss falls by at least 5x in 300 steps."""
    net, params, x, t, xg, tg, (n, h, w) = setup(2, n=4, h=16, w=16)
    wch = net.upload(np.ones(3))
    losses = []
    for it in range(300):
        net.forward(xg, n, h, w)
        losses.append(float(net.backward(tg, wch)))
        net.step(lr=3e-3)
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
