"""Density-fitted (T) integrals built on the GPU, with their adjoint.

Mirrors the (T) part of ``_make_df_eris_incore`` (pyscfad/cc/dfccsd.py:62-89):
from the MO-basis DF factors ``Lpq`` (naux, nmo, nmo) it forms

* ``ovvv = Lov^T . pack_tril(Lvv)`` -> (o, v, v(v+1)/2), packed like the reference,
* ``ovoo = Lov^T . Loo`` -> (o, v, o, o),
* ``ovov = Lov^T . Lov`` -> (o, v, o, v),

with ``pamd_df_eris`` (FP64 MFMA contractions), and the cotangent of ``Lpq``
from those of the three blocks with ``pamd_df_eris_vjp`` (the adjoints of
np.dot and of pack_tril, pyscfad/lib/_numpy_helper_opt.py:70-95).  The
integrals then never leave HBM: ``make_df_eris`` returns a ``ChemistsERIs``
whose blocks are autograd tensors, so ``ccsd_t.kernel`` on it chains dE/d(ovvv,
ovoo, ovov) into dE/dLpq.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ._ops import _dev, _ptr, _require_gpu, _stream
from .eris import ChemistsERIs


def _shape(Lpq, nocc):
    if Lpq.dim() != 3 or Lpq.shape[1] != Lpq.shape[2]:
        raise ValueError(f'Lpq must be (naux, nmo, nmo), got {tuple(Lpq.shape)}')
    naux, nmo = int(Lpq.shape[0]), int(Lpq.shape[1])
    if not 0 < nocc < nmo:
        raise ValueError(f'nocc={nocc} outside (0, nmo={nmo})')
    return naux, nmo, nmo - nocc


def _workspace(nocc, nmo, naux, with_vjp, device):
    lib = _lib.load()
    n = ctypes.c_size_t(0)
    _lib.check(lib.pamd_df_eris_workspace_size(nocc, nmo, naux, int(with_vjp), ctypes.byref(n)),
               'pamd_df_eris_workspace_size')
    return torch.empty(max(n.value, 1), dtype=torch.uint8, device=device), n.value


def df_eris(Lpq, nocc, blocks=('ovvv', 'ovoo', 'ovov')):
    """{name: device tensor} for the requested (T) blocks."""
    _require_gpu()
    lib = _lib.load()
    dev = Lpq.device if isinstance(Lpq, torch.Tensor) and Lpq.is_cuda else torch.device('cuda')
    L = _dev(Lpq, dev)
    naux, nmo, v = _shape(L, nocc)
    o = nocc
    shapes = dict(ovvv=(o, v, v * (v + 1) // 2), ovoo=(o, v, o, o), ovov=(o, v, o, v))
    out = {k: torch.empty(shapes[k], dtype=torch.float64, device=dev) for k in blocks}
    ws, nb = _workspace(o, nmo, naux, False, dev)
    st = lib.pamd_df_eris(o, nmo, naux, _ptr(L), _ptr(out.get('ovvv')), _ptr(out.get('ovoo')),
                          _ptr(out.get('ovov')), _ptr(ws), nb, _stream(dev))
    _lib.check(st, 'pamd_df_eris')
    return out


def df_eris_vjp(Lpq, nocc, ovvv_bar=None, ovoo_bar=None, ovov_bar=None):
    """Cotangent of Lpq (naux, nmo, nmo); a missing block cotangent counts as zero."""
    _require_gpu()
    lib = _lib.load()
    dev = Lpq.device if isinstance(Lpq, torch.Tensor) and Lpq.is_cuda else torch.device('cuda')
    L = _dev(Lpq, dev)
    naux, nmo, _v = _shape(L, nocc)
    gb = [_dev(x, dev) for x in (ovvv_bar, ovoo_bar, ovov_bar)]
    L_bar = torch.empty_like(L)
    ws, nb = _workspace(nocc, nmo, naux, True, dev)
    st = lib.pamd_df_eris_vjp(nocc, nmo, naux, _ptr(L), *[_ptr(x) for x in gb], _ptr(L_bar),
                              _ptr(ws), nb, _stream(dev))
    _lib.check(st, 'pamd_df_eris_vjp')
    return L_bar


class DFEris(torch.autograd.Function):
    """(ovvv packed, ovoo, ovov) = f(Lpq) with the HIP adjoint as backward."""

    @staticmethod
    def forward(ctx, Lpq, nocc):
        out = df_eris(Lpq, nocc)
        ctx.save_for_backward(Lpq)
        ctx.nocc = nocc
        return out['ovvv'], out['ovoo'], out['ovov']

    @staticmethod
    def backward(ctx, ovvv_bar, ovoo_bar, ovov_bar):
        (Lpq,) = ctx.saved_tensors
        return df_eris_vjp(Lpq, ctx.nocc, ovvv_bar, ovoo_bar, ovov_bar), None


def make_df_eris(Lpq, nocc, mo_energy, fock=None):
    """ChemistsERIs with GPU-built, differentiable (T) blocks (packed ovvv)."""
    L = Lpq if isinstance(Lpq, torch.Tensor) else _dev(Lpq, torch.device('cuda'))
    ovvv, ovoo, ovov = DFEris.apply(L, int(nocc))
    if fock is None:
        fock = torch.diag(mo_energy)
    return ChemistsERIs(fock=fock, mo_energy=mo_energy, ovoo=ovoo, ovov=ovov, ovvv=ovvv,
                        nocc=int(nocc))
