"""The allocation-free weight gradient (nemo2d_nn_conv_bwd_weight_noalloc)
against the entry point with a caller's work buffer, through the C-ABI: two
calls in a row, the second on a smaller batch, bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16_bits(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def operands(rng, n, h, w, cin, cout):
    from nemo2d.closure import DevArray, pad8
    x = np.zeros((n * h * w, pad8(cin)), np.uint16)
    x[:, :cin] = bf16_bits(rng.standard_normal((n * h * w, cin)))
    dz = np.zeros((n * h * w, pad8(cout)), np.uint16)
    dz[:, :cout] = bf16_bits(0.1 * rng.standard_normal((n * h * w, cout)))
    return DevArray.from_numpy(x.view(np.int16)), DevArray.from_numpy(dz.view(np.int16))


def test_noalloc_weight_gradient_equals_workspace_entry_point():
    import nemo2d
    from nemo2d.closure import DevArray, _setup_nn
    L = nemo2d.lib()
    _setup_nn(L)
    rng = np.random.default_rng(5)
    work = DevArray((int(L.nemo2d_nn_wgrad_work_floats(1)),), np.float32)
    for n, h, w, cin, cout in ((4, 96, 131, 32, 32), (3, 37, 50, 4, 32), (1, 9, 7, 32, 3)):
        print("noalloc weight gradient: batch %d x %d x %d, %d -> %d channels" % (n, h, w, cin, cout),
              flush=True)
        x, dz = operands(rng, n, h, w, cin, cout)
        dw_a, db_a = DevArray((9, cin, cout), np.float32), DevArray((cout,), np.float32)
        dw_b, db_b = DevArray((9, cin, cout), np.float32), DevArray((cout,), np.float32)
        rc = L.nemo2d_nn_conv_bwd_weight(x.ptr, dz.ptr, n, h, w, cin, cout, dw_a.fptr(), db_a.fptr(),
                                         work.fptr(), None)
        assert rc == 0, L.nemo2d_last_error()
        rc = L.nemo2d_nn_conv_bwd_weight_noalloc(x.ptr, dz.ptr, n, h, w, cin, cout, dw_b.fptr(),
                                                 db_b.fptr(), None)
        assert rc == 0, L.nemo2d_last_error()
        a, b = dw_a.download(), dw_b.download()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, h, w, np.abs(a - b).max())
        assert np.array_equal(db_a.download().view(np.uint32), db_b.download().view(np.uint32))
        assert np.abs(a).max() > 0


def test_noalloc_weight_gradient_rejects_too_many_pixels():
    import nemo2d
    from nemo2d.closure import DevArray, _setup_nn
    L = nemo2d.lib()
    _setup_nn(L)
    d = DevArray((16,), np.float32)
    rc = L.nemo2d_nn_conv_bwd_weight_noalloc(d.ptr, d.ptr, 1 << 12, 1 << 10, 1 << 10, 32, 32, d.fptr(),
                                             d.fptr(), None)
    assert rc == nemo2d.ERR_ARG and b"maximum" in L.nemo2d_last_error()
