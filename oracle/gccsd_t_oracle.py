"""Spin-orbital (T) from DF factors -- TEST INFRASTRUCTURE (cross-check only).

This module is test infrastructure (the oracle), never the product: only ``tests/`` may import it.

An independent statement of the (T) energy in the spin-orbital basis, from the
same spatial DF factors the closed-shell path starts from, used to cross-check
the closed-shell oracle (``ccsd_t_oracle``) and the GPU kernel at tiny sizes.
It follows the reference's GCCSD(T) with an occupied projector:

* spin orbitals interleave alpha/beta within the occupied and within the
  virtual block (orbspin = [0,1]*nocc + [0,1]*nvir), so spin orbital p is
  spatial orbital p // 2 with spin p % 2 -- pyscfad/cc/gccsd_t.py:98-109;
* RHF amplitudes mapped to spin orbitals by ``spatial2spin``
  (gccsd_t.py:24-84): t2aa = t2 - t2.transpose(1,0,2,3) on the same-spin
  blocks, t2 on (a_alpha b_beta | i_alpha j_beta), and the exchange-signed
  copies on the mixed blocks;
* antisymmetrised integrals <pq||rs> = (pr|qs) - (ps|qr) from the spin-blocked
  DF factors (gccsd_t.py:128-158);
* t3c, t3d, the antisymmetrising permutations, the denominators, the projector
  on the first occupied index and E = sum t3* w / 36 (gccsd_t.py:161-215).

O((2o)^3 (2v)^3) memory: tiny sizes only (o <= 4, v <= 10).
"""
from __future__ import annotations

import numpy as np


def spin_t1(t1):
    o, v = t1.shape
    s = np.zeros((2 * o, 2 * v))
    s[0::2, 0::2] = t1
    s[1::2, 1::2] = t1
    return s


def spin_t2(t2):
    o, _, v, _ = t2.shape
    aa = t2 - t2.transpose(1, 0, 2, 3)
    s = np.zeros((2 * o, 2 * o, 2 * v, 2 * v))
    s[0::2, 0::2, 0::2, 0::2] = aa
    s[1::2, 1::2, 1::2, 1::2] = aa
    s[0::2, 1::2, 0::2, 1::2] += t2                          # (i_a j_b | a_a b_b)
    s[1::2, 0::2, 1::2, 0::2] += t2.transpose(1, 0, 3, 2)    # (j_b i_a | b_b a_a)
    s[0::2, 1::2, 1::2, 0::2] -= t2.transpose(0, 1, 3, 2)    # (i_a j_b | b_b a_a)
    s[1::2, 0::2, 0::2, 1::2] -= t2.transpose(1, 0, 2, 3)    # (j_b i_a | a_a b_b)
    return s


def spin_eris(Lpq, nocc):
    """<pq||rs> over spin orbitals from spatial Lpq (naux, n, n)."""
    naux, n, _ = Lpq.shape
    spatial = np.arange(2 * n) // 2
    spin = np.arange(2 * n) % 2
    same = (spin[:, None] == spin[None, :]).astype(float)
    Ls = Lpq[:, spatial[:, None], spatial[None, :]] * same          # (naux, 2n, 2n)
    eri = np.einsum('Ppq,Prs->pqrs', Ls, Ls)                         # (pq|rs)
    return eri.transpose(0, 2, 1, 3) - eri.transpose(0, 2, 3, 1)      # <pq||rs>


def _asym_abc(x):
    x = x - x.transpose(0, 1, 2, 4, 3, 5) - x.transpose(0, 1, 2, 5, 4, 3)
    return x - x.transpose(1, 0, 2, 3, 4, 5) - x.transpose(2, 1, 0, 3, 4, 5)


def energy(Lpq, mo_energy, fock, t1, t2, nocc, prj=None):
    """Spin-orbital (T); prj (nfrag, nocc) projects the first occupied index
    (None = identity, the full E(T))."""
    o = nocc
    v = t1.shape[1]
    O = 2 * o
    g = spin_eris(Lpq, o)
    s1, s2 = spin_t1(t1), spin_t2(t2)
    n = o + v
    spatial = np.arange(2 * n) // 2
    spin = np.arange(2 * n) % 2
    fs = fock[spatial[:, None], spatial[None, :]] * (spin[:, None] == spin[None, :])
    e_s = np.asarray(mo_energy)[spatial]
    if prj is None:
        prj = np.eye(o)
    p1 = np.zeros((prj.shape[0], O))
    p1[:, 0::2] = prj
    p1[:, 1::2] = prj
    bcei = g[:O, O:, O:, O:].transpose(3, 2, 1, 0)
    majk = g[:O, :O, :O, O:].transpose(2, 3, 0, 1)
    bcjk = g[:O, :O, O:, O:].transpose(2, 3, 0, 1)
    eia = e_s[:O, None] - e_s[None, O:]
    d3 = (eia[:, None, None, :, None, None] + eia[None, :, None, None, :, None]
          + eia[None, None, :, None, None, :])
    t3c = (np.einsum('jkae,bcei->ijkabc', s2, bcei) - np.einsum('imbc,majk->ijkabc', s2, majk))
    t3c = _asym_abc(t3c)
    t3d = np.einsum('ia,bcjk->ijkabc', s1, bcjk)
    t3d += np.einsum('ai,jkbc->ijkabc', fs[O:, :O], s2)
    t3d = _asym_abc(t3d)
    t3 = (t3c + t3d) / d3
    t3 = np.einsum('pjkabc,ip->ijkabc', t3, p1)
    w = np.einsum('pjkabc,ip->ijkabc', t3c, p1)
    return np.einsum('ijkabc,ijkabc', t3, w) / 36
