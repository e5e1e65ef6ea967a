"""Host-side checks of every contraction the kernels will run (no GPU needed).

The library records the GEMM descriptors of one occupied triple; the emulator
evaluates them with bounds and write-race checks.  Numerics: the W build must
reproduce the oracle's W_ijk, and the adjoint contractions must equal torch
autograd of the oracle's W_ijk formula for a random cotangent Wbar.
"""
import numpy as np
import pytest
import torch

from pyscfad_amd.gen import make_problem
from oracle import ccsd_t_oracle as O
from tests.gemm_emulator import Memory, record, split_factors

CASES = [(3, 12, 0, 1, 2), (3, 12, 0, 0, 2), (3, 12, 1, 2, 2), (4, 17, 0, 2, 3), (2, 20, 0, 1, 1),
         (5, 9, 1, 1, 3)]


def _fill(mem, p):
    o, v = p.nocc, p.nvir
    mem.arrays['Vp'][:] = p.ovvv.transpose(0, 1, 3, 2).reshape(-1)          # Vp[x,p,q,f]
    mem.arrays['T2m'][:] = p.t2.transpose(0, 2, 3, 1).reshape(-1)          # T2m[x,p,q,m]
    mem.arrays['OOt'][:] = p.ovoo.transpose(0, 3, 1, 2).reshape(-1)        # OOt[x,y,p,m]
    mem.arrays['t2'][:] = p.t2.reshape(-1)


def _w_ijk_torch(ovvv, t2, ovoo, i, j, k):
    w = torch.einsum('iafb,kjcf->ijkabc', ovvv, t2) - torch.einsum('iamj,kmcb->ijkabc', ovoo, t2)
    W = sum(w.permute(*pm) for pm in O._PERMS)
    return W[i, j, k]


@pytest.mark.parametrize('o,v,i,j,k', CASES)
def test_forward_descriptors(o, v, i, j, k):
    p = make_problem(o, v, 3 * v, seed=o * 100 + v)
    mem = Memory(o, v)
    _fill(mem, p)
    descs = record(o, v, i, j, k, with_vjp=False)
    assert len(descs) == 3
    for d in descs:
        mem.run(d)
    W = O.p6(O._w_raw(p.t2, p.ovvv, p.ovoo))[i, j, k]
    got = mem.arrays['W'].reshape(v, v, v)
    assert np.abs(got - W).max() <= 1e-13 * max(1.0, np.abs(W).max())


@pytest.mark.parametrize('o,v,i,j,k', CASES)
def test_adjoint_descriptors(o, v, i, j, k):
    p = make_problem(o, v, 3 * v, seed=o * 100 + v + 1)
    mem = Memory(o, v)
    _fill(mem, p)
    rng = np.random.default_rng(7)
    wbar = rng.normal(size=(v, v, v))
    descs = record(o, v, i, j, k, with_vjp=True)
    assert len(descs) == 3 + 3 + 6 + 6 + 3
    for d in descs[:3]:
        mem.run(d)
    mem.arrays['W'][:] = wbar.reshape(-1)   # the epilogue overwrites W with Wbar
    for d in descs[3:]:
        mem.run(d)
    # torch autograd of the oracle's W_ijk formula
    ovvv = torch.tensor(p.ovvv, requires_grad=True)
    t2 = torch.tensor(p.t2, requires_grad=True)
    ovoo = torch.tensor(p.ovoo, requires_grad=True)
    Wt = _w_ijk_torch(ovvv, t2, ovoo, i, j, k)
    Wt.backward(torch.tensor(wbar))
    g_ovvv = ovvv.grad.numpy()
    g_t2 = t2.grad.numpy()
    g_ovoo = ovoo.grad.numpy()
    # ovvv_bar: Vpb[x,p,q,f] = dE/d ovvv[x,p,f,q]
    vpb = mem.arrays['Vpb'].reshape(o, v, v, v).transpose(0, 1, 3, 2)
    scale = max(1.0, np.abs(g_ovvv).max())
    assert np.abs(vpb - g_ovvv).max() <= 1e-12 * scale
    # t2_bar = hole-term T2mb[x,p,q,m] + particle partial slabs
    S_p, S_h = split_factors(o, v)
    t2b = mem.arrays['T2mb'].reshape(o, v, v, o).transpose(0, 3, 1, 2).copy()
    pp = mem.arrays['part_p'].reshape(6, S_p, v, v).sum(axis=1)
    for pr, (x, y) in enumerate([(k, j), (j, k), (k, i), (i, k), (j, i), (i, j)]):
        t2b[x, y] += pp[pr]
    assert np.abs(t2b - g_t2).max() <= 1e-12 * max(1.0, np.abs(g_t2).max())
    # ovoo_bar[x,p,m,y] from the hole partial slabs
    ph = mem.arrays['part_h'].reshape(6, S_h, v, o).sum(axis=1)
    oob = np.zeros((o, v, o, o))
    for pr, (x, y) in enumerate([(i, j), (i, k), (j, i), (j, k), (k, i), (k, j)]):
        oob[x, :, :, y] += ph[pr]
    assert np.abs(oob - g_ovoo).max() <= 1e-12 * max(1.0, np.abs(g_ovoo).max())


def _split_factors_prec(o, v, prec):
    import ctypes
    from pyscfad_amd import _lib
    lib = _lib.load()
    lib.pamd_debug_split_factors_prec.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)] * 2
    a, b = ctypes.c_int(), ctypes.c_int()
    assert lib.pamd_debug_split_factors_prec(o, v, prec, ctypes.byref(a), ctypes.byref(b)) == 0
    return a.value, b.value


@pytest.mark.parametrize('o,v', [(40, 400), (60, 600), (40, 800), (24, 240), (8, 40)])
def test_bf16x3_split_k_bounds_fp32_sum_length(o, v):
    """Every split-K part of the split-BF16x3 adjoint (K = v^2 terms) spans at
    most 8192 terms, so the FP32 tile sums of one part stay short at any size;
    and at most half the number of K stages of 32."""
    PREC_BF16X3 = 1
    S_p, S_h = _split_factors_prec(o, v, PREC_BF16X3)
    K = v * v
    for S in (S_p, S_h):
        assert 1 <= S <= max(1, K // 32)
        assert -(-K // S) <= 8192 or S == max(1, K // 32)
