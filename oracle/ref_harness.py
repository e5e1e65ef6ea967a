"""ctypes harness around the reference's own (T) C code -- TEST INFRASTRUCTURE.

``oracle/Makefile`` compiles pyscfadlib/pyscfadlib/vjp/cc/*.c (+ vjp/util) from
the reference tree into ``oracle/_ref/libcc_vjp_ref.so``.  This module restates
the few lines of numpy host glue the reference wraps around those C entry
points, so the C can run on synthetic inputs without jax or pyscf:

* layouts ``t1T``, ``t2T = t2.transpose(2,3,1,0)``, ``vooo``, ``vvop``
  -- pyscfad/cc/ccsd_t.py:88-106 and pyscfad/lno/ccsd_t.py:76-88;
* forward E(T) via ``lnoccsdt_contract`` over one block (0, nvir) with
  ``et = 2/3 e_tot`` -- pyscfad/lno/ccsd_t.py:90-112 (with ``mat = I`` this is
  the canonical E(T); with ``mat = u u^T`` the occupied-weighted one);
* VJP via ``ccsd_t_energy_vjp`` as one block ``contract(0, v, 0, v)`` with all
  four caches = ``vvop`` -- pyscfad/cc/ccsd_t.py:108-177 (a0 = 0 makes the
  b-loop empty), ``et_bar *= 2`` (line 86), output transforms lines 179-191;
* the local-orbital projected VJP via ``lnoccsdt_energy_vjp`` as one block
  with every cache = ``vvop`` -- pyscfad/lno/ccsd_t.py:124-237.

Only tests, the golden-vector script and bench.py's cpu_baseline use it.
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_ref', 'libcc_vjp_ref.so')
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def load():
    global _lib
    if _lib is None:
        # MKL dlopens its CPU-specific kernels, which resolve symbols from
        # libmkl_core: the chain must be loaded with global visibility.
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    return _lib


def _p(x):
    return x.ctypes.data_as(ctypes.c_void_p)


def _layouts(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo):
    nocc, nvir = t1.shape
    nmo = nocc + nvir
    mo_energy = np.ascontiguousarray(mo_energy, dtype=np.float64)
    t1T = np.ascontiguousarray(t1.T)
    t2T = np.ascontiguousarray(t2.transpose(2, 3, 1, 0))
    vooo = np.ascontiguousarray(ovoo.transpose(1, 0, 3, 2))
    fvo = np.ascontiguousarray(fvo)
    vvop = np.empty((nvir, nvir, nocc, nmo))
    vvop[:, :, :, :nocc] = ovov.transpose(1, 3, 0, 2)
    vvop[:, :, :, nocc:] = ovvv.transpose(1, 3, 0, 2)
    return mo_energy, t1T, t2T, vooo, fvo, vvop


def energy(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, mat=None):
    """E(T) from ``lnoccsdt_contract`` (one block), ``ovvv`` full (o,v,v,v)."""
    lib = load()
    nocc, nvir = t1.shape
    mo_energy, t1T, t2T, vooo, fvo, vvop = _layouts(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    if mat is None:
        mat = np.eye(nocc)
    mat = np.ascontiguousarray(mat, dtype=np.float64)
    e = np.zeros(1)
    lib.lnoccsdt_contract(_p(e), _p(mat), _p(mo_energy), _p(t1T), _p(t2T), _p(vooo), _p(fvo),
                          ctypes.c_int(nocc), ctypes.c_int(nvir), ctypes.c_int(0),
                          ctypes.c_int(nvir), _p(vvop))
    return float(e[0] * 2.0 / 3.0)


def energy_vjp(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, et_bar=1.0):
    """Reference cotangents in the reference's output layout.

    Returns a dict like ``ccsd_t_oracle.energy_vjp`` plus ``ovvv_bar_packed``
    (the reference's own fold) and ``ovvv_bar`` (unsymmetrised full block).
    """
    lib = load()
    nocc, nvir = t1.shape
    nmo = nocc + nvir
    mo_energy, t1T, t2T, vooo, fvo, vvop = _layouts(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    et_bar = 2.0 * et_bar
    mo_bar = np.zeros_like(mo_energy)
    t1T_bar = np.zeros_like(t1T)
    t2T_bar = np.zeros_like(t2T)
    vooo_bar = np.zeros_like(vooo)
    fvo_bar = np.zeros_like(fvo)
    vvop_bar = np.zeros_like(vvop)
    lib.ccsd_t_energy_vjp(
        _p(mo_energy), _p(t1T), _p(t2T), _p(vooo), _p(fvo), ctypes.c_double(et_bar),
        ctypes.c_int(nocc), ctypes.c_int(nvir),
        ctypes.c_int(0), ctypes.c_int(nvir), ctypes.c_int(0), ctypes.c_int(nvir),
        _p(vvop), _p(vvop), _p(vvop), _p(vvop),
        _p(mo_bar), _p(t1T_bar), _p(t2T_bar), _p(vooo_bar), _p(fvo_bar),
        _p(vvop_bar), _p(vvop_bar), _p(vvop_bar), _p(vvop_bar))
    t1_bar = np.ascontiguousarray(t1T_bar.T)
    t2_bar = np.ascontiguousarray(t2T_bar.transpose(3, 2, 0, 1))
    ovoo_bar = np.ascontiguousarray(vooo_bar.transpose(1, 0, 3, 2))
    ovov_bar = np.ascontiguousarray(vvop_bar[:, :, :, :nocc].transpose(2, 0, 3, 1))
    ovvv_full = np.ascontiguousarray(vvop_bar[:, :, :, nocc:].transpose(2, 0, 3, 1))
    g = ovvv_full + ovvv_full.transpose(0, 1, 3, 2)
    idx, idy = np.diag_indices(nvir)
    g[:, :, idx, idy] *= .5
    idx, idy = np.tril_indices(nvir)
    packed = np.ascontiguousarray(g[:, :, idx, idy])
    return dict(t1_bar=t1_bar, t2_bar=t2_bar, fvo_bar=fvo_bar, mo_energy_bar=mo_bar,
                ovoo_bar=ovoo_bar, ovov_bar=ovov_bar, ovvv_bar=ovvv_full,
                ovvv_bar_packed=packed)


def projected_energy_vjp(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, ulo, et_bar=1.0):
    """Occupied-weighted (T) of mat = ulo^T ulo and its VJP from the reference C.

    Forward: ``lnoccsdt_contract`` with that mat (``energy(..., mat=...)``).
    Backward: ``lnoccsdt_energy_vjp`` as one block, b = c = (0, nvir), every
    cache = ``vvop`` (pyscfad/lno/ccsd_t.py:205-214, first iteration of its
    block loop with bufsize >= nvir), ``et_bar *= 2/3`` (line 132) and the
    output transforms of lines 226-234.  ``ovvv`` is full (o,v,v,v); the packed
    cotangent is the reference's own fold.  Returns (E, dict) with the blocks
    of ``energy_vjp`` plus ``mat_bar`` (o,o) and ``ulo_bar = ulo (mat_bar +
    mat_bar^T)``, the chain rule of mat = ulo^T ulo (pyscfad/lno/ccsd_t.py:42).
    """
    lib = load()
    nocc, nvir = t1.shape
    ulo = np.ascontiguousarray(np.atleast_2d(ulo), dtype=np.float64)
    mat = np.ascontiguousarray(ulo.T @ ulo)
    e = energy(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo, mat=mat)
    mo_energy, t1T, t2T, vooo, fvo, vvop = _layouts(mo_energy, t1, t2, ovvv, ovoo, ovov, fvo)
    eb = et_bar * 2.0 / 3.0
    mat_bar = np.zeros_like(mat)
    mo_bar = np.zeros_like(mo_energy)
    t1T_bar = np.zeros_like(t1T)
    t2T_bar = np.zeros_like(t2T)
    vooo_bar = np.zeros_like(vooo)
    fvo_bar = np.zeros_like(fvo)
    vvop_bar = np.zeros_like(vvop)
    lib.lnoccsdt_energy_vjp(
        _p(mat), _p(mo_energy), _p(t1T), _p(t2T), _p(vooo), _p(fvo), ctypes.c_double(eb),
        ctypes.c_int(nocc), ctypes.c_int(nvir),
        ctypes.c_int(0), ctypes.c_int(nvir), ctypes.c_int(0), ctypes.c_int(nvir),
        _p(vvop), _p(vvop), _p(vvop), _p(vvop),
        _p(mat_bar), _p(mo_bar), _p(t1T_bar), _p(t2T_bar), _p(vooo_bar), _p(fvo_bar),
        _p(vvop_bar), _p(vvop_bar), _p(vvop_bar), _p(vvop_bar))
    ovvv_full = np.ascontiguousarray(vvop_bar[:, :, :, nocc:].transpose(2, 0, 3, 1))
    g = ovvv_full + ovvv_full.transpose(0, 1, 3, 2)
    idx, idy = np.diag_indices(nvir)
    g[:, :, idx, idy] *= .5
    idx, idy = np.tril_indices(nvir)
    out = dict(t1_bar=np.ascontiguousarray(t1T_bar.T),
               t2_bar=np.ascontiguousarray(t2T_bar.transpose(3, 2, 0, 1)),
               fvo_bar=fvo_bar, mo_energy_bar=mo_bar,
               ovoo_bar=np.ascontiguousarray(vooo_bar.transpose(1, 0, 3, 2)),
               ovov_bar=np.ascontiguousarray(vvop_bar[:, :, :, :nocc].transpose(2, 0, 3, 1)),
               ovvv_bar=ovvv_full, ovvv_bar_packed=np.ascontiguousarray(g[:, :, idx, idy]),
               mat_bar=mat_bar, ulo_bar=ulo @ (mat_bar + mat_bar.T))
    return e, out
