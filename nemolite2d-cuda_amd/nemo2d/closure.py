"""Learned closure for a 2x-coarsened NEMOLite2D run (no counterpart in the
reference).

* DevArray -- a typed device buffer owned by the library
  (nemo2d_dev_alloc/free/upload/download/zero); the network needs no other
  device runtime.
* ConvNet -- 4 'same' 3x3 convolutions (4 -> 32 -> 32 -> 32 -> 3 channels),
  exact GELU between them, BF16 weights/activations with FP32 accumulation,
  trained with Adam on FP32 master weights.  Every operation runs in the
  library's own kernels (libnemo2d.so, nemo2d_nn_*: BF16-MFMA implicit-GEMM
  convolutions forward and backward, GELU backward, MSE, Adam) on the
  default stream.
* Closure data generation, training and evaluation live in
  closure_pipeline (same package).
"""
import ctypes

import numpy as np

from . import _check, lib

LAYERS = ((4, 32, True), (32, 32, True), (32, 32, True), (32, 3, False))


def pad8(c):
    return (c + 7) // 8 * 8


class DevArray:
    """Device buffer of `shape` and numpy `dtype` (float32 or int16 = BF16
    bits) on the current HIP device, zero-initialised."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._L = lib()
        p = ctypes.c_void_p()
        _check(self._L.nemo2d_dev_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p
        self.zero()

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def size(self):
        return int(np.prod(self.shape))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("DevArray.upload: %d bytes into %d" % (a.nbytes, self.nbytes))
        _check(self._L.nemo2d_dev_upload(self.ptr, a.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self._L.nemo2d_dev_download(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        _check(self._L.nemo2d_dev_zero(self.ptr, self.nbytes))

    def fptr(self):
        return ctypes.cast(self.ptr, ctypes.POINTER(ctypes.c_float))

    def close(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self._L.nemo2d_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _setup_nn(L):
    if getattr(L, "_nn_ready", False):
        return
    vp, i, l, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    L.nemo2d_nn_conv_fwd.argtypes = [vp, i, i, i, i, vp, fp, i, fp, vp, vp, i, vp]
    L.nemo2d_nn_conv_bwd_data.argtypes = [vp, i, i, i, i, vp, i, fp, vp]
    L.nemo2d_nn_conv_bwd_data_gelu.argtypes = [vp, i, i, i, i, vp, i, vp, i, vp, vp]
    L.nemo2d_nn_conv_bwd_weight.argtypes = [vp, vp, i, i, i, i, i, fp, fp, fp, vp]
    L.nemo2d_nn_conv_bwd_weight_noalloc.argtypes = [vp, vp, i, i, i, i, i, fp, fp, vp]
    L.nemo2d_nn_wgrad_work_floats.argtypes = [l]
    L.nemo2d_dev_synchronize.argtypes = []
    L.nemo2d_nn_wgrad_work_floats.restype = l
    L.nemo2d_nn_gelu_bwd.argtypes = [fp, fp, l, vp, vp]
    L.nemo2d_nn_mse.argtypes = [fp, fp, l, i, fp, vp, fp, i, vp]
    L.nemo2d_nn_adam.argtypes = [fp, fp, fp, fp, l, f, f, f, f, i, vp]
    L.nemo2d_nn_pack_weights.argtypes = [fp, i, i, vp, vp, vp]
    L.nemo2d_nn_to_bf16.argtypes = [fp, l, vp, vp]
    for n in ("conv_fwd", "conv_bwd_data", "conv_bwd_data_gelu", "conv_bwd_weight",
              "conv_bwd_weight_noalloc", "gelu_bwd", "mse", "adam", "pack_weights", "to_bf16"):
        getattr(L, "nemo2d_nn_" + n).restype = i
    L._nn_ready = True


class ConvNet:
    """The closure network on the current HIP device."""

    # weights start N(0, init_scale^2 * 2 / (9 cin)); the default scale
    # sqrt(1/6) gives variance 1 / (27 cin), that of the uniform fan-in rule
    # U(-1/sqrt(9 cin), 1/sqrt(9 cin)).  The He scale (1.0) trains to a lower
    # loss but the coarse+closure run can then drift away from the fine one.
    INIT_SCALE = float(np.sqrt(1.0 / 6.0))
    STREAM = None  # the default stream

    def __init__(self, seed=0, layers=LAYERS, init_scale=None, fp32_preact=False):
        """fp32_preact: keep the hidden layers' pre-activations (read by the
        GELU backward) in FP32 instead of BF16; the output layer's is FP32
        either way."""
        self.fp32_preact = fp32_preact
        self.L = lib()
        _setup_nn(self.L)
        self.layers = layers
        rng = np.random.default_rng(seed)
        if init_scale is None:
            init_scale = self.INIT_SCALE
        self.w, self.b, self.wf, self.wb = [], [], [], []
        self.mw, self.vw, self.mb, self.vb = [], [], [], []
        for cin, cout, _ in layers:
            std = init_scale * np.sqrt(2.0 / (9 * cin))
            self.w.append(DevArray.from_numpy(rng.standard_normal((9, cin, cout)) * std, np.float32))
            self.b.append(DevArray((cout,), np.float32))
            self.wf.append(DevArray((9 * 32 * pad8(cin),), np.int16))
            self.wb.append(DevArray((9 * 32 * pad8(cout),), np.int16))
            for lst, shape in ((self.mw, (9, cin, cout)), (self.vw, (9, cin, cout)),
                               (self.mb, (cout,)), (self.vb, (cout,))):
                lst.append(DevArray(shape, np.float32))
        self.step_count = 0
        self.pack()
        self._shape = None
        self._sets = {}

    def pack(self):
        for k, (cin, cout, _) in enumerate(self.layers):
            _check(self.L.nemo2d_nn_pack_weights(self.w[k].fptr(), cin, cout, self.wf[k].ptr,
                                                 self.wb[k].ptr, self.STREAM))

    def set_weights(self, ws, bs):
        """FP32 numpy weights [9, cin, cout] and biases [cout]."""
        for k in range(len(self.layers)):
            self.w[k].upload(np.asarray(ws[k], dtype=np.float32))
            self.b[k].upload(np.asarray(bs[k], dtype=np.float32))
        self.pack()

    def weights(self):
        return ([w.download().astype(np.float64) for w in self.w],
                [b.download().astype(np.float64) for b in self.b])

    def upload(self, a):
        """FP32 device copy of a host array (targets, channel weights)."""
        return DevArray.from_numpy(a, np.float32)

    def to_bf16(self, x):
        """FP32 host array [P, C] (C a multiple of 8) -> device BF16 bits."""
        xf = DevArray.from_numpy(x, np.float32)
        out = DevArray(xf.shape, np.int16)
        _check(self.L.nemo2d_nn_to_bf16(xf.fptr(), xf.size, out.ptr, self.STREAM))
        return out

    def _buffers(self, n, h, w):
        """Activation and gradient buffers for batches of n images of h x w
        (kept per shape, so that alternating shapes do not reallocate)."""
        if self._shape == (n, h, w):
            return
        if self._shape is not None:
            self._sets[self._shape] = self._bufset()
        if (n, h, w) in self._sets:
            self._use(self._sets.pop((n, h, w)))
            self._shape = (n, h, w)
            return
        P = n * h * w
        nl = len(self.layers)
        self.z = [DevArray((P, pad8(co)), np.float32 if (k == nl - 1 or self.fp32_preact) else np.int16)
                  for k, (_, co, _) in enumerate(self.layers)]
        self.act = [DevArray((P, pad8(co)), np.int16) for _, co, _ in self.layers]
        self.dz = [DevArray((P, pad8(co)), np.int16) for _, co, _ in self.layers]
        if not hasattr(self, "gw"):  # gradients and reduction space: one set for all shapes
            self.gw = [DevArray((9, ci, co), np.float32) for ci, co, _ in self.layers]
            self.gb = [DevArray((co,), np.float32) for _, co, _ in self.layers]
            self.work = DevArray((int(self.L.nemo2d_nn_wgrad_work_floats(P)),), np.float32)
            self.nparts = 512
            self.part = DevArray((self.nparts,), np.float32)
        self._shape = (n, h, w)

    _BUF_KEYS = ("z", "act", "dz")

    def _bufset(self):
        return {k: getattr(self, k) for k in self._BUF_KEYS}

    def _use(self, d):
        for k, v in d.items():
            setattr(self, k, v)

    def forward(self, x_bf16, n, h, w, train=True):
        """x_bf16: device [n*h*w, 8] BF16 bits (channels 0..3 used).  Returns
        the FP32 output, a device [P, 8] array (channels 0..2).  train=False
        (inference) writes no hidden pre-activations: they are read only by
        backward."""
        self._buffers(n, h, w)
        self.x = x_bf16
        a = x_bf16
        nl = len(self.layers)
        for k, (cin, cout, g) in enumerate(self.layers):
            last = k == nl - 1
            f32 = self.z[k].dtype == np.float32
            keep = train or last
            _check(self.L.nemo2d_nn_conv_fwd(a.ptr, n, h, w, cin, self.wf[k].ptr, self.b[k].fptr(),
                                             cout, self.z[k].fptr() if f32 and keep else None,
                                             None if f32 or not keep else self.z[k].ptr,
                                             None if last else self.act[k].ptr, 1 if g else 0,
                                             self.STREAM))
            a = self.act[k]
        return self.z[-1]

    def backward(self, target, wch, want_loss=True):
        """target: device FP32 [P, 8]; wch: device FP32 [3] channel weights.
        Gradients into self.gw/self.gb; returns the loss (None when
        want_loss is false: no host synchronisation)."""
        n, h, w = self._shape
        P = n * h * w
        nl = len(self.layers)
        _check(self.L.nemo2d_nn_mse(self.z[-1].fptr(), target.fptr(), P, self.layers[-1][1],
                                    wch.fptr(), self.dz[-1].ptr, self.part.fptr(), self.nparts,
                                    self.STREAM))
        for k in range(nl - 1, -1, -1):
            cin, cout, _ = self.layers[k]
            xin = self.act[k - 1] if k > 0 else self.x
            _check(self.L.nemo2d_nn_conv_bwd_weight(xin.ptr, self.dz[k].ptr, n, h, w, cin, cout,
                                                    self.gw[k].fptr(), self.gb[k].fptr(),
                                                    self.work.fptr(), self.STREAM))
            if k > 0:
                # the layer below's dz = dX * GELU'(z below), in one pass
                zb = self.z[k - 1].dtype != np.float32
                _check(self.L.nemo2d_nn_conv_bwd_data_gelu(self.dz[k].ptr, n, h, w, cout,
                                                           self.wb[k].ptr, cin, self.z[k - 1].ptr,
                                                           1 if zb else 0, self.dz[k - 1].ptr,
                                                           self.STREAM))
        if not want_loss:
            return None
        return float(self.part.download().astype(np.float64).sum())

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for k in range(len(self.layers)):
            for p, g, m, v in ((self.w[k], self.gw[k], self.mw[k], self.vw[k]),
                               (self.b[k], self.gb[k], self.mb[k], self.vb[k])):
                _check(self.L.nemo2d_nn_adam(p.fptr(), g.fptr(), m.fptr(), v.fptr(), p.size, lr, b1,
                                             b2, eps, self.step_count, self.STREAM))
        self.pack()

    def time_kernels(self, x_bf16, n, h, w, reps=20, layer=1):
        """Seconds per launch of the convolution kernels of one 32 -> 32
        layer (forward as in training, forward as in inference (activations
        only), data gradient with GELU backward, weight gradient with the
        partial sum) on a batch of n images h x w, each launched reps times
        back to back after a forward/backward pass (host clock around the
        launches, device synchronised on both sides)."""
        import time
        self.forward(x_bf16, n, h, w)
        cin, cout, _ = self.layers[layer]
        k = layer
        fw = lambda: self.L.nemo2d_nn_conv_fwd(self.act[k - 1].ptr, n, h, w, cin, self.wf[k].ptr,  # noqa: E731
                                               self.b[k].fptr(), cout, None, self.z[k].ptr,
                                               self.act[k].ptr, 1, self.STREAM)
        fi = lambda: self.L.nemo2d_nn_conv_fwd(self.act[k - 1].ptr, n, h, w, cin, self.wf[k].ptr,  # noqa: E731
                                               self.b[k].fptr(), cout, None, None,
                                               self.act[k].ptr, 1, self.STREAM)
        bd = lambda: self.L.nemo2d_nn_conv_bwd_data_gelu(self.dz[k].ptr, n, h, w, cout,  # noqa: E731
                                                         self.wb[k].ptr, cin, self.z[k - 1].ptr, 1,
                                                         self.dz[k - 1].ptr, self.STREAM)
        bw = lambda: self.L.nemo2d_nn_conv_bwd_weight(self.act[k - 1].ptr, self.dz[k].ptr, n, h, w,  # noqa: E731
                                                      cin, cout, self.gw[k].fptr(), self.gb[k].fptr(),
                                                      self.work.fptr(), self.STREAM)
        out = {}
        for name, f in (("conv_fwd", fw), ("conv_fwd_inference", fi), ("conv_bwd_data_gelu", bd),
                        ("conv_bwd_weight", bw)):
            _check(f())
            _check(self.L.nemo2d_dev_synchronize())
            t0 = time.perf_counter()
            for _ in range(reps):
                _check(f())
            _check(self.L.nemo2d_dev_synchronize())
            out[name] = (time.perf_counter() - t0) / reps
        return out

    def grads(self):
        return ([g.download().astype(np.float64) for g in self.gw],
                [g.download().astype(np.float64) for g in self.gb])
