"""Drop-in fragment (T) of a local-orbital projector, with its adjoint.

Mirrors pyscfad/lno/ccsd_t.py:35-60 -- ``kernel(mycc, eris, ulo, t1=None,
t2=None, verbose=logger.NOTE)`` returns the occupied-weighted E(T) with
``mat = ulo^T ulo`` and logs ``CCSD(T) correction = %.15g``.  With torch
inputs it is a ``torch.autograd.Function``: the forward runs
``pamd_ccsdt_projected_energy`` and the backward ``pamd_ccsdt_projected_energy_vjp``,
returning cotangents for ``ulo, mo_energy, fock, t1, t2, ovvv, ovoo, ovov``
(``fock_bar`` is zero outside its vo block, as in pyscfad/cc/ccsd_t.py:179-191).
Precision: ``mycc.ccsd_t_precision`` or env ``PYSCFAD_AMD_CCSD_T_PRECISION``.
"""
from __future__ import annotations

import numpy as np

from ..cc import _ops
from ..cc.ccsd_t import _precision
from ..lib import logger


def kernel(mycc, eris, ulo, t1=None, t2=None, verbose=logger.NOTE, precision=None):
    import torch
    if t1 is None:
        t1 = mycc.t1
    if t2 is None:
        t2 = mycc.t2
    prec = _precision(mycc, precision)
    log = logger.new_logger(mycc, verbose)

    class _ProjectedT(torch.autograd.Function):
        @staticmethod
        def forward(ctx, ulo_, mo_energy, fock, t1_, t2_, ovvv, ovoo, ovov):
            nocc = t1_.shape[0]
            inp = _ops.Inputs(mo_energy, t1_, t2_, ovvv, ovoo, ovov, fock[nocc:, :nocc],
                              precision=prec)
            ctx.inp = inp
            ctx.ulo = ulo_.detach()
            ctx.dev_in = [x.device for x in (ulo_, mo_energy, fock, t1_, t2_, ovvv, ovoo, ovov)]
            return _ops.projected_energy(inp, ctx.ulo).to(t1_.device)

        @staticmethod
        def backward(ctx, ybar):
            inp = ctx.inp
            _e, g = _ops.projected_energy_vjp(inp, ctx.ulo, et_bar=float(ybar.detach().cpu()))
            o, v = inp.nocc, inp.nvir
            fock_bar = torch.zeros((o + v, o + v), dtype=torch.float64, device=inp.device)
            fock_bar[o:, :o] = g['fvo_bar']
            out = (g['ulo_bar'], g['mo_energy_bar'], fock_bar, g['t1_bar'], g['t2_bar'],
                   g['ovvv_bar'], g['ovoo_bar'], g['ovov_bar'])
            return tuple(x.to(d) for x, d in zip(out, ctx.dev_in))

    def as_t(x):
        if isinstance(x, torch.Tensor):
            return x
        return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64)))

    u = as_t(ulo)
    if u.dim() == 1:
        u = u[None, :]
    et = _ProjectedT.apply(u, as_t(eris.mo_energy), as_t(eris.fock), as_t(t1), as_t(t2),
                           as_t(eris.ovvv), as_t(eris.ovoo), as_t(eris.ovov))
    log.note('CCSD(T) correction = %.15g', float(et.detach().cpu()))
    return et
