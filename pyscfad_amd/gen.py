"""Seeded synthetic inputs for the closed-shell (T) energy and its adjoint.

The generator builds a density-fitting-like problem from a seed:

* ``L[P,p,q]`` (naux, n, n) with entries N(0, sigma = 0.1/sqrt(naux)), symmetric
  in (p, q): the upper triangle is drawn and mirrored.
* occupied orbital energies U[-1.5, -0.3], virtual ones U[0.1, 3.0], each sorted
  ascending; ``fock = diag(mo_energy)`` so ``fvo = 0`` unless ``fvo_scale`` is set.
* ERI blocks in chemists' notation, as pyscfad stores them
  (pyscfad/cc/dfccsd.py:62-89, pyscfad/cc/ccsd.py:30-31):
  ``ovvv[i,a,b,c] = (ia|bc)``, ``ovoo[i,a,j,k] = (ia|jk)``, ``ovov[i,a,j,b] = (ia|jb)``.
* MP2 amplitudes ``t2[i,j,a,b] = ovov[i,a,j,b] / (e_i + e_j - e_a - e_b)``
  (pyscfad/cc/ccsd.py:376-379) and ``t1 ~ N(0, 1e-3)``.

Draw order from ``numpy.random.default_rng(seed)``: the (naux, n(n+1)/2) upper
triangles of L, then e_occ, e_vir, t1, then (only if ``fvo_scale``) fvo.

Small problems are built on the host in float64 with numpy.  Large ones
(``device`` given) draw the same numbers on the host and form the ERI blocks on
the GPU with torch float64 matmuls, so a problem built on either side is the
same problem up to rounding of the contractions.
"""
from __future__ import annotations

import dataclasses
import numpy as np


@dataclasses.dataclass
class Problem:
    nocc: int
    nvir: int
    naux: int
    mo_energy: object      # (n,)
    fock: object           # (n, n)
    t1: object             # (o, v)
    t2: object             # (o, o, v, v)
    ovvv: object           # (o, v, v, v) full
    ovoo: object           # (o, v, o, o)
    ovov: object           # (o, v, o, v)

    @property
    def fvo(self):
        return self.fock[self.nocc:, :self.nocc]

    def ovvv_packed(self):
        """(o, v, v(v+1)/2) lower-triangle packing of the last two indices."""
        v = self.nvir
        idx, idy = np.tril_indices(v)
        return self.ovvv[:, :, idx, idy]


def _draw(nocc, nvir, naux, seed, fvo_scale=0.0):
    n = nocc + nvir
    rng = np.random.default_rng(seed)
    sigma = 0.1 / np.sqrt(naux)
    iu, ju = np.triu_indices(n)
    tri = rng.normal(0.0, sigma, size=(naux, iu.size))
    e_occ = np.sort(rng.uniform(-1.5, -0.3, size=nocc))
    e_vir = np.sort(rng.uniform(0.1, 3.0, size=nvir))
    t1 = rng.normal(0.0, 1e-3, size=(nocc, nvir))
    fvo = None
    if fvo_scale:
        fvo = rng.normal(0.0, fvo_scale, size=(nvir, nocc))
    return (iu, ju, tri), e_occ, e_vir, t1, fvo


def _unpack_L(iu, ju, tri, naux, n, xp, **kw):
    L = xp.zeros((naux, n, n), dtype=xp.float64, **kw)
    L[:, iu, ju] = tri
    L[:, ju, iu] = tri
    return L


def make_problem(nocc, nvir, naux, seed=0, fvo_scale=0.0, device=None,
                 L_perturb=None, e_shift=None):
    """Build a :class:`Problem`.

    ``device`` None -> numpy float64 arrays; otherwise torch float64 tensors on
    that device (the ERI contractions run there).

    ``L_perturb``: optional ``(scale, seed)``: adds ``scale * N(0,1)`` (symmetric
    as L) to L, drawn from ``default_rng(seed)`` after the base draws.
    ``e_shift``: optional ``(sigma, seed)``: adds N(0, sigma) to every orbital
    energy (then re-sorts each block), drawn from ``default_rng(seed)``.
    """
    n = nocc + nvir
    (iu, ju, tri), e_occ, e_vir, t1, fvo = _draw(nocc, nvir, naux, seed, fvo_scale)
    if L_perturb is not None:
        scale, pseed = L_perturb
        prng = np.random.default_rng(pseed)
        tri = tri + scale * prng.normal(0.0, 1.0, size=tri.shape)
    if e_shift is not None:
        sig, eseed = e_shift
        erng = np.random.default_rng(eseed)
        e_occ = np.sort(e_occ + erng.normal(0.0, sig, size=nocc))
        e_vir = np.sort(e_vir + erng.normal(0.0, sig, size=nvir))
    mo_energy = np.concatenate([e_occ, e_vir])
    fock = np.diag(mo_energy)
    if fvo is not None:
        fock[nocc:, :nocc] = fvo
        fock[:nocc, nocc:] = fvo.T

    if device is None:
        L = _unpack_L(iu, ju, tri, naux, n, np)
        Lov = L[:, :nocc, nocc:].reshape(naux, nocc * nvir)
        Lvv = L[:, nocc:, nocc:].reshape(naux, nvir * nvir)
        Loo = L[:, :nocc, :nocc].reshape(naux, nocc * nocc)
        ovvv = (Lov.T @ Lvv).reshape(nocc, nvir, nvir, nvir)
        ovoo = (Lov.T @ Loo).reshape(nocc, nvir, nocc, nocc)
        ovov = (Lov.T @ Lov).reshape(nocc, nvir, nocc, nvir)
        eia = e_occ[:, None] - e_vir[None, :]
        t2 = ovov.transpose(0, 2, 1, 3) / (eia[:, None, :, None] + eia[None, :, None, :])
        return Problem(nocc, nvir, naux, mo_energy, fock, t1, np.ascontiguousarray(t2),
                       ovvv, ovoo, ovov)

    import torch
    dev = torch.device(device)
    tri_t = torch.from_numpy(tri).to(dev)
    L = unpack_L_torch(torch.from_numpy(iu).to(dev), torch.from_numpy(ju).to(dev), tri_t, naux, n)
    del tri_t
    return assemble_torch(L, mo_energy, fock, t1, nocc, nvir)


def make_L(nocc, nvir, naux, seed=0, device=None):
    """The symmetric DF factors L (naux, n, n) that ``make_problem`` draws for
    the same arguments (numpy, or torch on ``device``)."""
    (iu, ju, tri), *_ = _draw(nocc, nvir, naux, seed)
    n = nocc + nvir
    if device is None:
        return _unpack_L(iu, ju, tri, naux, n, np)
    import torch
    dev = torch.device(device)
    return unpack_L_torch(torch.from_numpy(iu).to(dev), torch.from_numpy(ju).to(dev),
                          torch.from_numpy(tri).to(dev), naux, n)


def unpack_L_torch(iu_t, ju_t, tri_t, naux, n):
    import torch
    L = torch.zeros((naux, n, n), dtype=torch.float64, device=tri_t.device)
    L[:, iu_t, ju_t] = tri_t
    L[:, ju_t, iu_t] = tri_t
    return L


def assemble_torch(L, mo_energy, fock, t1, nocc, nvir):
    """ERI blocks and MP2 t2 on L's device from a symmetric L (naux, n, n)."""
    import torch
    dev = L.device
    naux = L.shape[0]
    Lov = L[:, :nocc, nocc:].reshape(naux, nocc * nvir)
    Loo = L[:, :nocc, :nocc].reshape(naux, nocc * nocc)
    # ovvv in slabs of the occupied index to bound the temporary
    ovvv = torch.empty((nocc, nvir, nvir, nvir), dtype=torch.float64, device=dev)
    Lvv = L[:, nocc:, nocc:].reshape(naux, nvir * nvir).contiguous()
    for i in range(nocc):
        Li = L[:, i, nocc:]
        ovvv[i] = (Li.T @ Lvv).reshape(nvir, nvir, nvir)
    del Lvv
    ovoo = (Lov.T @ Loo).reshape(nocc, nvir, nocc, nocc).contiguous()
    ovov = (Lov.T @ Lov).reshape(nocc, nvir, nocc, nvir).contiguous()
    del Lov, Loo
    me = torch.as_tensor(mo_energy, dtype=torch.float64, device=dev)
    eia = me[:nocc, None] - me[None, nocc:]
    t2 = (ovov.permute(0, 2, 1, 3) /
          (eia[:, None, :, None] + eia[None, :, None, :])).contiguous()
    return Problem(nocc, nvir, naux, me,
                   torch.as_tensor(fock, dtype=torch.float64, device=dev),
                   torch.as_tensor(t1, dtype=torch.float64, device=dev),
                   t2, ovvv, ovoo, ovov)
