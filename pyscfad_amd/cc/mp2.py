"""MP2 amplitudes and occupied-pair MP2 energies on the GPU.

``init_amps`` mirrors the t2/emp2 part of ``CCSD.init_amps``
(pyscfad/cc/ccsd.py:364-381): ``t2[i,j,a,b] = ovov[i,a,j,b] / (eia[i,a] +
eia[j,b])`` with ``eia = e_occ[:, None] - e_vir[None, :]``, and the MP2 energy
``2 sum t2[i,j,a,b] ovov[i,a,j,b] - sum t2[j,i,a,b] ovov[i,a,j,b]``, here
resolved over ordered occupied pairs (i, j) by ``pamd_mp2_init_amps``.
"""
from __future__ import annotations

import torch

from .. import _lib
from ._ops import _dev, _ptr, _require_gpu, _stream


def init_amps(mo_energy, ovov, nocc, with_t2=True):
    """(emp2, t2 or None, e_pair (o, o)) as float64 device tensors."""
    _require_gpu()
    lib = _lib.load()
    dev = ovov.device if isinstance(ovov, torch.Tensor) and ovov.is_cuda else torch.device('cuda')
    g = _dev(ovov, dev)
    e = _dev(mo_energy, dev)
    o = int(nocc)
    if g.dim() != 4 or g.shape[0] != o or g.shape[2] != o or g.shape[1] != g.shape[3]:
        raise ValueError(f'ovov must be (o, v, o, v) with o={o}, got {tuple(g.shape)}')
    v = int(g.shape[1])
    if e.shape != (o + v,):
        raise ValueError(f'mo_energy must have {o + v} entries, got {tuple(e.shape)}')
    t2 = torch.empty((o, o, v, v), dtype=torch.float64, device=dev) if with_t2 else None
    e_pair = torch.empty((o, o), dtype=torch.float64, device=dev)
    st = lib.pamd_mp2_init_amps(o, v, _ptr(e), _ptr(g), _ptr(t2), _ptr(e_pair), _stream(dev))
    _lib.check(st, 'pamd_mp2_init_amps')
    return e_pair.sum(), t2, e_pair


def pair_energies(mo_energy, ovov, nocc):
    """(o, o) MP2 pair energies of the MP2 amplitudes (no t2 written)."""
    return init_amps(mo_energy, ovov, nocc, with_t2=False)[2]
