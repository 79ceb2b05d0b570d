"""Learned closure for a 2x-coarsened NEMOLite2D run (no counterpart in the
reference).

* ConvNet -- 4 'same' 3x3 convolutions (4 -> 32 -> 32 -> 32 -> 3 channels),
  exact GELU between them, BF16 weights/activations with FP32 accumulation,
  trained with Adam on FP32 master weights.  Every operation runs in the
  library's own kernels (libnemo2d.so, nemo2d_nn_*: BF16-MFMA implicit-GEMM
  convolutions forward and backward, GELU backward, MSE, Adam); torch only
  allocates device memory and supplies the stream.
* restrict_state / features / apply_correction -- the fine->coarse
  restriction on the C-grid (volume-conserving ssh average, depth-weighted
  face velocities), the network's input channels and the correction.
* Closure data generation, training and evaluation live in
  closure_pipeline (same package).
"""
import ctypes

import numpy as np

from . import _check, lib

LAYERS = ((4, 32, True), (32, 32, True), (32, 32, True), (32, 3, False))


def pad8(c):
    return (c + 7) // 8 * 8


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _fptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_float))


def _setup_nn(L):
    if getattr(L, "_nn_ready", False):
        return
    vp, i, l, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    L.nemo2d_nn_conv_fwd.argtypes = [vp, i, i, i, i, vp, fp, i, fp, vp, i, vp]
    L.nemo2d_nn_conv_bwd_data.argtypes = [vp, i, i, i, i, vp, i, fp, vp]
    L.nemo2d_nn_conv_bwd_weight.argtypes = [vp, vp, i, i, i, i, i, fp, fp, fp, vp]
    L.nemo2d_nn_wgrad_work_floats.argtypes = [l]
    L.nemo2d_nn_wgrad_work_floats.restype = l
    L.nemo2d_nn_gelu_bwd.argtypes = [fp, fp, l, vp, vp]
    L.nemo2d_nn_mse.argtypes = [fp, fp, l, i, fp, vp, fp, i, vp]
    L.nemo2d_nn_adam.argtypes = [fp, fp, fp, fp, l, f, f, f, f, i, vp]
    L.nemo2d_nn_pack_weights.argtypes = [fp, i, i, vp, vp, vp]
    L.nemo2d_nn_to_bf16.argtypes = [fp, l, vp, vp]
    for n in ("conv_fwd", "conv_bwd_data", "conv_bwd_weight", "gelu_bwd", "mse", "adam",
              "pack_weights", "to_bf16"):
        getattr(L, "nemo2d_nn_" + n).restype = i
    L._nn_ready = True


class ConvNet:
    """The closure network on the current HIP device."""

    def __init__(self, seed=0, layers=LAYERS, init_scale=1.0):
        import torch
        self.torch = torch
        self.L = lib()
        _setup_nn(self.L)
        self.layers = layers
        rng = np.random.default_rng(seed)
        dev = "cuda"
        self.w, self.b, self.wf, self.wb = [], [], [], []
        self.mw, self.vw, self.mb, self.vb = [], [], [], []
        for cin, cout, _ in layers:
            std = init_scale * np.sqrt(2.0 / (9 * cin))
            w = torch.tensor(rng.standard_normal((9, cin, cout)) * std, dtype=torch.float32, device=dev)
            self.w.append(w.contiguous())
            self.b.append(torch.zeros(cout, dtype=torch.float32, device=dev))
            self.wf.append(torch.zeros(9 * 32 * pad8(cin), dtype=torch.int16, device=dev))
            self.wb.append(torch.zeros(9 * 32 * pad8(cout), dtype=torch.int16, device=dev))
            for lst, shape in ((self.mw, (9, cin, cout)), (self.vw, (9, cin, cout)),
                               (self.mb, (cout,)), (self.vb, (cout,))):
                lst.append(torch.zeros(shape, dtype=torch.float32, device=dev))
        self.step_count = 0
        self.pack()
        self._shape = None

    @property
    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def pack(self):
        for k, (cin, cout, _) in enumerate(self.layers):
            _check(self.L.nemo2d_nn_pack_weights(_fptr(self.w[k]), cin, cout, _ptr(self.wf[k]),
                                                 _ptr(self.wb[k]), self.stream))

    def set_weights(self, ws, bs):
        """FP32 numpy weights [9, cin, cout] and biases [cout]."""
        for k in range(len(self.layers)):
            self.w[k].copy_(self.torch.tensor(ws[k], dtype=self.torch.float32))
            self.b[k].copy_(self.torch.tensor(bs[k], dtype=self.torch.float32))
        self.pack()

    def weights(self):
        return ([w.cpu().numpy().astype(np.float64) for w in self.w],
                [b.cpu().numpy().astype(np.float64) for b in self.b])

    def to_bf16(self, x):
        """FP32 [P, C] tensor (C a multiple of 8) -> BF16 bits (int16)."""
        out = self.torch.empty(x.shape, dtype=self.torch.int16, device=x.device)
        _check(self.L.nemo2d_nn_to_bf16(_fptr(x), x.numel(), _ptr(out), self.stream))
        return out

    def _buffers(self, n, h, w):
        if self._shape == (n, h, w):
            return
        t = self.torch
        P = n * h * w
        self.z = [t.zeros((P, pad8(co)), dtype=t.float32, device="cuda") for _, co, _ in self.layers]
        self.act = [t.zeros((P, pad8(co)), dtype=t.int16, device="cuda") for _, co, _ in self.layers]
        self.dz = [t.zeros((P, pad8(co)), dtype=t.int16, device="cuda") for _, co, _ in self.layers]
        self.dact = t.zeros((P, 32), dtype=t.float32, device="cuda")
        self.gw = [t.zeros((9, ci, co), dtype=t.float32, device="cuda") for ci, co, _ in self.layers]
        self.gb = [t.zeros(co, dtype=t.float32, device="cuda") for _, co, _ in self.layers]
        self.work = t.empty(int(self.L.nemo2d_nn_wgrad_work_floats(P)), dtype=t.float32, device="cuda")
        self.nparts = 512
        self.part = t.zeros(self.nparts, dtype=t.float32, device="cuda")
        self._shape = (n, h, w)

    def forward(self, x_bf16, n, h, w):
        """x_bf16: [n*h*w, 8] BF16 bits (channels 0..3 used).  Returns the
        FP32 output [P, 8] (channels 0..2)."""
        self._buffers(n, h, w)
        self.x = x_bf16
        a = x_bf16
        for k, (cin, cout, g) in enumerate(self.layers):
            _check(self.L.nemo2d_nn_conv_fwd(_ptr(a), n, h, w, cin, _ptr(self.wf[k]), _fptr(self.b[k]),
                                             cout, _fptr(self.z[k]), _ptr(self.act[k]), 1 if g else 0,
                                             self.stream))
            a = self.act[k]
        return self.z[-1]

    def backward(self, target, wch):
        """target: FP32 [P, 8]; wch: FP32 tensor [3] of channel weights.
        Gradients into self.gw/self.gb; returns the loss (a device tensor)."""
        n, h, w = self._shape
        P = n * h * w
        nl = len(self.layers)
        _check(self.L.nemo2d_nn_mse(_fptr(self.z[-1]), _fptr(target), P, self.layers[-1][1], _fptr(wch),
                                    _ptr(self.dz[-1]), _fptr(self.part), self.nparts, self.stream))
        for k in range(nl - 1, -1, -1):
            cin, cout, _ = self.layers[k]
            xin = self.act[k - 1] if k > 0 else self.x
            _check(self.L.nemo2d_nn_conv_bwd_weight(_ptr(xin), _ptr(self.dz[k]), n, h, w, cin, cout,
                                                    _fptr(self.gw[k]), _fptr(self.gb[k]),
                                                    _fptr(self.work), self.stream))
            if k > 0:
                dact = self.torch.as_strided(self.dact, (P, pad8(cin)), (pad8(cin), 1))
                _check(self.L.nemo2d_nn_conv_bwd_data(_ptr(self.dz[k]), n, h, w, cout, _ptr(self.wb[k]),
                                                      cin, _fptr(dact), self.stream))
                _check(self.L.nemo2d_nn_gelu_bwd(_fptr(dact), _fptr(self.z[k - 1]), P * pad8(cin),
                                                 _ptr(self.dz[k - 1]), self.stream))
        return self.part.sum()

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for k in range(len(self.layers)):
            for p, g, m, v in ((self.w[k], self.gw[k], self.mw[k], self.vw[k]),
                               (self.b[k], self.gb[k], self.mb[k], self.vb[k])):
                _check(self.L.nemo2d_nn_adam(_fptr(p), _fptr(g), _fptr(m), _fptr(v), p.numel(), lr, b1,
                                             b2, eps, self.step_count, self.stream))
        self.pack()

    def grads(self):
        return ([g.cpu().numpy().astype(np.float64) for g in self.gw],
                [g.cpu().numpy().astype(np.float64) for g in self.gb])
