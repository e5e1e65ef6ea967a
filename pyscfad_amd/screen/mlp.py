"""BF16 MLP trained and evaluated by the library's HIP kernels.

Layers: in -> 256 -> 256 -> out, SiLU between, bf16 weights and activations,
FP32 accumulation (``pamd_bf16_gemm_tn``), FP32 master weights updated by
``pamd_adam_step``.  torch only allocates device memory and orders the calls on
its current stream; all arithmetic runs in the kernels of include/pyscfad_amd_mlp.h.
"""
from __future__ import annotations

import ctypes
import math

import torch

from .. import _lib

_SIGS = {
    'pamd_bf16_gemm_tn': [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    'pamd_bf16_act_fwd': [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'pamd_bf16_act_bwd': [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'pamd_adam_step': [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6 +
                      [ctypes.c_float] * 5 + [ctypes.c_int, ctypes.c_void_p],
    'pamd_mse': [ctypes.c_int] + [ctypes.c_void_p] * 5,
    'pamd_f32_to_bf16': [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                         ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
}


def _fn(name):
    lib = _lib.load()
    f = getattr(lib, name)
    f.argtypes = _SIGS[name]
    f.restype = ctypes.c_int
    return f


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _s(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def gemm_tn(A, B, bias=None, out=None, accumulate=False):
    """C = A B^T (+ bias) with bf16 (int16-typed) A [M,K], B [N,K]; fp32 C [M,N]."""
    M, K = A.shape
    N = B.shape[0]
    if out is None:
        out = torch.empty((M, N), dtype=torch.float32, device=A.device)
    st = _fn('pamd_bf16_gemm_tn')(M, N, K, _p(A), A.stride(0), _p(B), B.stride(0), _p(bias),
                                  _p(out), out.stride(0), int(accumulate), _s(A.device))
    _lib.check(st, 'pamd_bf16_gemm_tn')
    return out


def to_bf16(x, with_t=False):
    """fp32 [M,N] -> bf16 bit patterns (int16 tensor) [M,N] (and [N,M])."""
    M, N = x.shape
    y = torch.empty((M, N), dtype=torch.int16, device=x.device)
    yt = torch.empty((N, M), dtype=torch.int16, device=x.device) if with_t else None
    st = _fn('pamd_f32_to_bf16')(M * N, _p(x.contiguous()), _p(y), M, N, _p(yt), _s(x.device))
    _lib.check(st, 'pamd_f32_to_bf16')
    return (y, yt) if with_t else y


def bf16_to_f32(b):
    """bf16 bit patterns (int16) -> float32 values (exact)."""
    return (b.to(torch.int32) & 0xFFFF).bitwise_left_shift(16).view(torch.float32)


class MLP:
    def __init__(self, sizes, device, seed=0):
        self.sizes = list(sizes)
        self.dev = torch.device(device)
        g = torch.Generator(device='cpu').manual_seed(seed)
        self.layers = []
        for fan_in, fan_out in zip(sizes[:-1], sizes[1:]):
            w = (torch.randn((fan_out, fan_in), generator=g) * math.sqrt(2.0 / fan_in)).float()
            b = torch.zeros(fan_out)
            L = dict(w=w.to(self.dev), b=b.to(self.dev))
            L['wb'], L['wt'] = to_bf16(L['w'], with_t=True)
            for k in ('w', 'b'):
                L['m_' + k] = torch.zeros_like(L[k])
                L['v_' + k] = torch.zeros_like(L[k])
            self.layers.append(L)
        self.step_count = 0

    def forward(self, X16, keep=False):
        """X16: bf16 [B, in]; returns fp32 [B, out]; keeps activations when ``keep``."""
        acts = [(X16, None)]
        h = X16
        n = len(self.layers)
        Z = None
        for li, L in enumerate(self.layers):
            Z = gemm_tn(h, L['wb'], bias=L['b'])
            if li < n - 1:
                B, N = Z.shape
                Y = torch.empty((B, N), dtype=torch.int16, device=self.dev)
                Yt = torch.empty((N, B), dtype=torch.int16, device=self.dev) if keep else None
                _lib.check(_fn('pamd_bf16_act_fwd')(B, N, _p(Z), 1, _p(Y), _p(Yt), _s(self.dev)),
                           'pamd_bf16_act_fwd')
                acts.append((Y, Yt, Z))
                h = Y
        if keep:
            self._acts = acts
        return Z

    def backward(self, dY):
        """Gradients of the last forward(keep=True) for output cotangent dY (fp32)."""
        acts = self._acts
        grads = [None] * len(self.layers)
        g = dY.contiguous()
        B = g.shape[0]
        for li in range(len(self.layers) - 1, -1, -1):
            L = self.layers[li]
            N, K = L['w'].shape
            act = 0 if li == len(self.layers) - 1 else 1
            Zl = acts[li + 1][2] if act else g  # pre-activation of this layer's output
            dZ = torch.empty((B, N), dtype=torch.int16, device=self.dev)
            dZt = torch.empty((N, B), dtype=torch.int16, device=self.dev)
            db = torch.empty(N, dtype=torch.float32, device=self.dev)
            _lib.check(_fn('pamd_bf16_act_bwd')(B, N, _p(Zl), _p(g), act, _p(dZ), _p(dZt), _p(db),
                                                _s(self.dev)), 'pamd_bf16_act_bwd')
            # input activations of this layer, transposed [K, B]
            if li == 0:
                Xt = self._x_t
            else:
                Xt = acts[li][1]
            dW = gemm_tn(dZt, Xt)                 # [N, K] = dZ^T X
            grads[li] = (dW, db)
            if li > 0:
                g = gemm_tn(dZ, L['wt'])          # [B, K] = dZ W
        return grads

    def adam(self, grads, lr=1e-3, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        f = _fn('pamd_adam_step')
        for L, (dW, db) in zip(self.layers, grads):
            N, K = L['w'].shape
            _lib.check(f(N, K, _p(L['w']), _p(dW), _p(L['m_w']), _p(L['v_w']), _p(L['wb']),
                         _p(L['wt']), lr, b1, b2, eps, wd, self.step_count, _s(self.dev)),
                       'pamd_adam_step')
            _lib.check(f(1, N, _p(L['b']), _p(db), _p(L['m_b']), _p(L['v_b']), None, None,
                         lr, b1, b2, eps, 0.0, self.step_count, _s(self.dev)), 'pamd_adam_step')

    def mse(self, pred, y, want_grad=True):
        n = pred.numel()
        loss = torch.empty((), dtype=torch.float32, device=self.dev)
        d = torch.empty_like(pred) if want_grad else None
        _lib.check(_fn('pamd_mse')(n, _p(pred), _p(y), _p(d), _p(loss), _s(self.dev)), 'pamd_mse')
        return loss, d

    def grads(self, X16, X16_t, y):
        """(prediction, loss, per-layer (dW, db)) of the MSE loss on one batch."""
        self._x_t = X16_t
        pred = self.forward(X16, keep=True)
        loss, d = self.mse(pred, y)
        return pred, loss, self.backward(d)

    def train_step(self, X16, X16_t, y, lr, wd=0.0):
        _, loss, g = self.grads(X16, X16_t, y)
        self.adam(g, lr=lr, wd=wd)
        return loss
