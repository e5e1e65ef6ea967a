"""CPU oracle for the closed-shell (T) energy and its VJP -- TEST INFRASTRUCTURE.

This module is test infrastructure (the oracle), never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it.  The product path (``pyscfad_amd``) never calls it and fails loudly when its
HIP library is missing.

It restates, in vectorised numpy over the full 6-index (i,j,k,a,b,c) space, the
algorithm of the reference's optimised (T) path with the reference's RAW index
conventions (which operand element each term reads), so that gradients match the
reference block by block:

* ``w_abc[i,j,k] = sum_f ovvv[i,a,f,b] t2[k,j,c,f] - sum_m ovoo[i,a,m,j] t2[k,m,c,b]``
  -- pyscfadlib/pyscfadlib/vjp/cc/ccsd_t.c:121-142 (``get_wv``: the two dgemm
  calls on ``t2T = t2.transpose(2,3,1,0)``, ``vv_op = vvop[a,b]`` and
  ``vooo = ovoo.transpose(1,0,3,2)``; layouts from pyscfad/cc/ccsd_t.py:88-106).
* ``v_abc[i,j,k] = 1/2 (ovov[i,a,j,b] t1[k,c] + t2[j,i,b,a] fvo[c,k])``
  -- ccsd_t.c:144-150.
* ``W = sum over the 6 joint permutations`` of (ijk, abc) -- ccsd_t.c:74-96
  (``_make_permute_indices``) with ccsd_t_vjp.c:32-37 (``get_wz``).
* ``Z = r3(W + V)``, r3(x)[ijk] = 4x[ijk]+x[jki]+x[kij]-2x[kji]-2x[ikj]-2x[jik]
  -- ccsd_t.c:98-119 (``add_and_permute``).
* ``E = 2 sum_{a>=b>=c} s_abc sum_ijk W Z / D = 1/3 sum_{abc} sum_ijk W Z / D``
  -- ccsd_t_vjp.c:298-304 (weights), pyscfad/cc/ccsd_t.py:86 (factor 2),
  pyscfad/cc/ccsd_t_slow.py:118-124, pyscfad/lno/ccsd_t.py:109-112.
* VJP: ccsd_t_vjp.c:55-269 (``get_w0v0d3_bar``, ``get_mo_energy_bar``,
  ``get_eris_amps_bar``); output transforms pyscfad/cc/ccsd_t.py:179-191
  (t1/t2 transposes, ``fock_bar`` vo block, packed-ovvv fold).

Memory is o^3 v^3 doubles per array, so this is meant for small problems
(o <= 8, v <= 40 finishes in seconds).

Pinned by: golden vectors from the reference's own C compiled under
``oracle/_ref`` (tests/golden), torch autograd of an independent brute-force
statement (``brute_energy``), closed forms (o=1 or v=1 gives E=0), exact
homogeneity / affinity identities and central finite differences (tests/).
"""
from __future__ import annotations

import numpy as np

# transposes that realise the 6 joint permutations on a (i,j,k,a,b,c) array:
# P[sigma](w)[i,j,k,a,b,c] = w[sigma(ijk), sigma(abc)]
_PERMS = (
    (0, 1, 2, 3, 4, 5),   # w_abc[ijk]
    (0, 2, 1, 3, 5, 4),   # w_acb[ikj]
    (1, 0, 2, 4, 3, 5),   # w_bac[jik]
    (2, 0, 1, 5, 3, 4),   # w_bca[jki]
    (1, 2, 0, 4, 5, 3),   # w_cab[kij]
    (2, 1, 0, 5, 4, 3),   # w_cba[kji]
)


def p6(w):
    """Joint-permutation symmetrisation; self-adjoint (sum over the group)."""
    out = np.zeros_like(w)
    for p in _PERMS:
        out += w.transpose(p)
    return out


def r3(x):
    """r3 over the occupied axes (0,1,2) at fixed (a,b,c); r3^T == r3."""
    return (4 * x + x.transpose(1, 2, 0, 3, 4, 5) + x.transpose(2, 0, 1, 3, 4, 5)
            - 2 * x.transpose(2, 1, 0, 3, 4, 5) - 2 * x.transpose(0, 2, 1, 3, 4, 5)
            - 2 * x.transpose(1, 0, 2, 3, 4, 5))


def _denom(mo_energy, nocc):
    eo = mo_energy[:nocc]
    ev = mo_energy[nocc:]
    return (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None]
            + eo[None, None, :, None, None, None] - ev[None, None, None, :, None, None]
            - ev[None, None, None, None, :, None] - ev[None, None, None, None, None, :])


def _w_raw(t2, ovvv, ovoo):
    w = np.einsum('iafb,kjcf->ijkabc', ovvv, t2, optimize=True)
    w -= np.einsum('iamj,kmcb->ijkabc', ovoo, t2, optimize=True)
    return w


def _v_raw(t1, t2, ovov, fvo):
    v = np.einsum('iajb,kc->ijkabc', ovov, t1, optimize=True)
    v += np.einsum('jiba,ck->ijkabc', t2, fvo, optimize=True)
    return 0.5 * v


def _forward(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo):
    nocc = t1.shape[0]
    W = p6(_w_raw(t2, ovvv, ovoo))
    V = p6(_v_raw(t1, t2, ovov, fvo))
    Z = r3(W + V)
    D = _denom(mo_energy, nocc)
    return W, V, Z, D


def energy(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo):
    """E(T); ``ovvv`` full (o,v,v,v), ``fvo`` = fock[o:, :o]."""
    W, V, Z, D = _forward(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    return float(np.sum(W * Z / D) / 3.0)


def pair_energy(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo):
    """Occupied-pair resolved E(T): e[i,j] = 1/3 sum_{k,abc} W Z / D; sum = E(T).

    Same weighted (T) the reference evaluates with an occupied projector
    (pyscfad/lno/ccsd_t.py:42, pyscfadlib/.../lno_ccsd_t.c:16-26): with
    ``mat = e_i e_i^T`` it returns sum_j e[i,j].
    """
    W, V, Z, D = _forward(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    return np.einsum('ijkabc->ij', W * Z / D) / 3.0


def projected_energy(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, ulo):
    """Occupied-weighted E(T) of the local-orbital projector mat = ulo^T ulo.

    pyscfad/lno/ccsd_t.py:42 forms ``mat = ulo.T @ ulo``; lno_ccsd_t.c:16-26
    contracts W over its first occupied index with mat before the energy
    product: E = 1/3 sum_{l,i,j,k,abc} mat[l,i] W[l,j,k,abc] Z[i,j,k,abc] / D[i,j,k,abc]
    (mat = I gives E(T); mat = e_i e_i^T gives the row sums of ``pair_energy``).
    """
    ulo = np.asarray(ulo, dtype=np.float64)
    mat = ulo.T @ ulo
    W, V, Z, D = _forward(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    return float(np.einsum('li,ljkabc,ijkabc->', mat, W, Z / D, optimize=True) / 3.0)


def energy_vjp(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, et_bar=1.0, occ_weight=None):
    """(E, cotangents) with ``et_bar`` the cotangent of E.

    ``occ_weight`` m (nocc,) gives the occupied-weighted energy of a diagonal
    projector, E_m = 1/3 sum m_i W Z / D = sum_i m_i sum_j pair_energy[i, j]
    (lnoccsdt_contract with mat = diag(m), lno_ccsd_t.c:16-26).  As
    sum_abc W Z / D is symmetric in (i, j, k), E_m also equals the sum with the
    symmetric weight mu_ijk = (m_i + m_j + m_k) / 3, and that form's seed
    mu / 3 replaces 1/3 in the adjoint below.

    Returns the raw-convention gradient blocks in the reference's output layout
    (pyscfad/cc/ccsd_t.py:179-191): ``t1_bar`` (o,v), ``t2_bar`` (o,o,v,v),
    ``fvo_bar`` (v,o), ``mo_energy_bar`` (n,), ``ovoo_bar`` (o,v,o,o),
    ``ovov_bar`` (o,v,o,v), ``ovvv_bar`` full (o,v,v,v) unsymmetrised
    (``fold_ovvv_packed`` gives the packed cotangent).
    """
    nocc, nvir = t1.shape
    W, V, Z, D = _forward(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    if occ_weight is None:
        e = float(np.sum(W * Z / D) / 3.0)
        g = et_bar / 3.0
    else:
        m = np.asarray(occ_weight, dtype=np.float64)
        mu = (m[:, None, None] + m[None, :, None] + m[None, None, :]) / 3.0
        mu = mu[:, :, :, None, None, None]
        e = float(np.sum(mu * W * Z / D) / 3.0)
        g = et_bar * mu / 3.0
    Y = W / D
    r3Y = r3(Y)
    Wbar = g * (Z / D + r3Y)
    Vbar = g * r3Y
    Dbar = -g * W * Z / D ** 2
    mo_bar, g = _backprop(nocc, t1, t2, ovvv, ovoo, ovov, fvo, Wbar, Vbar, Dbar)
    return e, dict(g, mo_energy_bar=mo_bar)


def _backprop(nocc, t1, t2, ovvv, ovoo, ovov, fvo, Wbar, Vbar, Dbar):
    """Operand cotangents from the seeds dE/dW, dE/dV, dE/dD of the full arrays
    (ccsd_t_vjp.c:98-269): D = e_i + e_j + e_k - e_a - e_b - e_c, W = P6(w),
    V = P6(v) with w, v the raw terms of ``_w_raw`` / ``_v_raw``."""
    nvir = t1.shape[1]
    mo_bar = np.zeros(nocc + nvir)
    mo_bar[:nocc] = (np.einsum('ijkabc->i', Dbar) + np.einsum('ijkabc->j', Dbar)
                     + np.einsum('ijkabc->k', Dbar))
    mo_bar[nocc:] = -(np.einsum('ijkabc->a', Dbar) + np.einsum('ijkabc->b', Dbar)
                      + np.einsum('ijkabc->c', Dbar))
    wbar = p6(Wbar)
    vbar = p6(Vbar)
    ovvv_bar = np.einsum('ijkabc,kjcf->iafb', wbar, t2, optimize=True)
    t2_bar = np.einsum('ijkabc,iafb->kjcf', wbar, ovvv, optimize=True)
    ovoo_bar = -np.einsum('ijkabc,kmcb->iamj', wbar, t2, optimize=True)
    t2_bar -= np.einsum('ijkabc,iamj->kmcb', wbar, ovoo, optimize=True)
    ovov_bar = 0.5 * np.einsum('ijkabc,kc->iajb', vbar, t1, optimize=True)
    t1_bar = 0.5 * np.einsum('ijkabc,iajb->kc', vbar, ovov, optimize=True)
    t2_bar += 0.5 * np.einsum('ijkabc,ck->jiba', vbar, fvo, optimize=True)
    fvo_bar = 0.5 * np.einsum('ijkabc,jiba->ck', vbar, t2, optimize=True)
    return mo_bar, dict(t1_bar=t1_bar, t2_bar=t2_bar, fvo_bar=fvo_bar, ovoo_bar=ovoo_bar,
                        ovov_bar=ovov_bar, ovvv_bar=ovvv_bar)


def projected_energy_vjp(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, ulo, et_bar=1.0):
    """(E, cotangents) of ``projected_energy``, plus ``ulo_bar``.

    With Y = Z / D: E = 1/3 sum mat[l,i] W[l,...] Y[i,...], so the seeds are
    dE/dW = P + r3(Q / D), dE/dV = r3(Q / D), dE/dD = -Q Z / D^2, with
    P = 1/3 sum_i mat[l,i] Y[i,...] and Q = 1/3 sum_l mat[l,i] W[l,...]
    (r3 is its own adjoint).  mat_bar[l,i] = 1/3 <W[l,...], Y[i,...]> and
    ulo_bar = ulo (mat_bar + mat_bar^T) for mat = ulo^T ulo.
    """
    nocc = t1.shape[0]
    ulo = np.asarray(ulo, dtype=np.float64)
    mat = ulo.T @ ulo
    W, V, Z, D = _forward(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    Y = Z / D
    e = float(np.einsum('li,ljkabc,ijkabc->', mat, W, Y, optimize=True) / 3.0)
    P = et_bar * np.einsum('li,ijkabc->ljkabc', mat, Y, optimize=True) / 3.0
    Q = et_bar * np.einsum('li,ljkabc->ijkabc', mat, W, optimize=True) / 3.0
    Xbar = r3(Q / D)
    mo_bar, g = _backprop(nocc, t1, t2, ovvv, ovoo, ovov, fvo, P + Xbar, Xbar, -Q * Z / D ** 2)
    mat_bar = et_bar * np.einsum('ljkabc,ijkabc->li', W, Y, optimize=True) / 3.0
    return e, dict(g, mo_energy_bar=mo_bar, mat_bar=mat_bar, ulo_bar=ulo @ (mat_bar + mat_bar.T))


def fold_ovvv_packed(ovvv_bar):
    """Cotangent of the packed (o,v,v(v+1)/2) ovvv from the full-array one.

    pyscfad/cc/ccsd_t.py:186-191: g + g^T over the last two axes, diagonal
    halved, lower triangle taken in ``tril_indices`` order.
    """
    nvir = ovvv_bar.shape[-1]
    g = ovvv_bar + ovvv_bar.transpose(0, 1, 3, 2)
    idx, idy = np.diag_indices(nvir)
    g[:, :, idx, idy] *= 0.5
    idx, idy = np.tril_indices(nvir)
    return np.ascontiguousarray(g[:, :, idx, idy])


def fock_bar_from_fvo(fvo_bar, nocc):
    nvir = fvo_bar.shape[0]
    fb = np.zeros((nocc + nvir, nocc + nvir))
    fb[nocc:, :nocc] = fvo_bar
    return fb


def brute_energy(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, xp=np):
    """Independent statement of E(T) by explicit loops over (a,b,c).

    Follows pyscfad/cc/ccsd_t_slow.py:74-121 (``get_w``/``get_v``/``r3``/
    ``body``) but with the C path's operand indices (ccsd_t.c:137-150), and
    sums all (a,b,c) with weight 1/3.  Works for numpy or torch (``xp``) so that
    torch autograd can differentiate it; meant for tiny sizes only.
    """
    nocc, nvir = t1.shape
    eo = mo_energy[:nocc]
    ev = mo_energy[nocc:]
    eijk = eo[:, None, None] + eo[None, :, None] + eo[None, None, :]
    es = 'einsum'

    def ein(sub, *ops):
        return getattr(xp, es)(sub, *ops)

    def get_w(a, b, c):
        w = ein('if,kjf->ijk', ovvv[:, a, :, b], t2[:, :, c, :])
        w = w - ein('imj,km->ijk', ovoo[:, a, :, :], t2[:, :, c, b])
        return w

    def get_v(a, b, c):
        v = ein('ij,k->ijk', ovov[:, a, :, b], t1[:, c])
        v = v + ein('ji,k->ijk', t2[:, :, b, a], fvo[c, :])
        return 0.5 * v

    def r3l(x):
        return (4 * x + x.permute(1, 2, 0) + x.permute(2, 0, 1) - 2 * x.permute(2, 1, 0)
                - 2 * x.permute(0, 2, 1) - 2 * x.permute(1, 0, 2)) if xp is not np else \
            (4 * x + x.transpose(1, 2, 0) + x.transpose(2, 0, 1) - 2 * x.transpose(2, 1, 0)
             - 2 * x.transpose(0, 2, 1) - 2 * x.transpose(1, 0, 2))

    def tp(x, p):
        return x.permute(*p) if xp is not np else x.transpose(p)

    et = 0.0
    for a in range(nvir):
        for b in range(nvir):
            for c in range(nvir):
                W = (get_w(a, b, c) + tp(get_w(a, c, b), (0, 2, 1))
                     + tp(get_w(b, a, c), (1, 0, 2)) + tp(get_w(b, c, a), (2, 0, 1))
                     + tp(get_w(c, a, b), (1, 2, 0)) + tp(get_w(c, b, a), (2, 1, 0)))
                V = (get_v(a, b, c) + tp(get_v(a, c, b), (0, 2, 1))
                     + tp(get_v(b, a, c), (1, 0, 2)) + tp(get_v(b, c, a), (2, 0, 1))
                     + tp(get_v(c, a, b), (1, 2, 0)) + tp(get_v(c, b, a), (2, 1, 0)))
                d3 = eijk - ev[a] - ev[b] - ev[c]
                et = et + (W * r3l(W + V) / d3).sum()
    return et / 3.0
