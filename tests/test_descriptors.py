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
from tests.gemm_emulator import NAMES, Memory, record, record_paired, sizes, split_factors

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


def test_direct_engine_eligibility():
    """Which contractions run on the FP64 engine's direct-to-LDS path: every
    W-build and adjoint contraction at nocc=40/nvir=400 and at the golden-vector
    sizes (even sizes, 16-byte aligned slabs); not every one when nocc and nvir
    are odd (element pairs are then not 16-byte aligned)."""
    import ctypes
    from pyscfad_amd import _lib
    lib = _lib.load()
    for o, v in ((40, 400), (8, 40), (4, 12), (6, 20)):
        descs = record(o, v, 0, 1, 2, with_vjp=True)
        ok = [lib.pamd_debug_gemm_direct_ok(ctypes.byref(d)) for d in descs]
        assert ok == [1] * len(descs), (o, v, ok)
    descs = record(3, 17, 0, 1, 2, with_vjp=True)
    assert 0 in [lib.pamd_debug_gemm_direct_ok(ctypes.byref(d)) for d in descs]


def test_bf16x3_direct_engine_eligibility():
    """Which contractions run on the split-BF16x3 engine's direct-to-LDS path
    (16-byte runs of 8 bf16 along each operand's fast index): every W-build and
    adjoint contraction when nocc and nvir are multiples of 8 (the benchmark
    and the adjoint-check sizes); when nocc is not (10, 60), the contractions
    with hole terms stay on the register-staged path (row strides that are
    multiples of nocc are not 16-byte aligned), while the ovvv and particle
    t2 adjoint contractions still qualify."""
    import ctypes
    from pyscfad_amd import _lib
    lib = _lib.load()

    def ok(o, v):
        return [lib.pamd_debug_gemm_b3p_ok(ctypes.byref(d), o * v ** 3)
                for d in record(o, v, 0, 1, 2, with_vjp=True)]
    for o, v in ((40, 400), (24, 240), (16, 160), (8, 40)):
        assert ok(o, v) == [1] * len(ok(o, v)), (o, v)
    # descriptor order: W build (3), ovvv adjoint (3), particle t2 (6), hole ovoo (6), hole t2 (3)
    for o, v in ((10, 400), (60, 600)):
        assert ok(o, v) == [0] * 3 + [1] * 9 + [0] * 9, (o, v)
    assert 0 in ok(6, 20)
    # a plane stride that breaks 16-byte alignment of the mid and lo planes
    d = record(40, 400, 0, 1, 2, with_vjp=False)[0]
    assert lib.pamd_debug_gemm_b3p_ok(ctypes.byref(d), 40 * 400 ** 3 + 4) == 0


def _rfast_swz(R, k):
    """Chunk XOR of k-row k in a row-contiguous tile of R rows (gemm_direct.hip)."""
    return ((k >> 2) & 3) << 3 if R % 64 == 0 else ((k >> 2) & 3) << 1


def _rfast80_at(k, r):
    """LDS offset (doubles) of row r, k-row k in an 80-row row-contiguous stage:
    the 16-row blocks of k-rows with (k >> 2) even take the even 128-byte block
    slots, the others the odd ones (gemm_direct.hip rfast_at)."""
    kr = (k & 3) + 4 * (k >> 3)
    bslot = 2 * (5 * kr + (r >> 4)) + ((k >> 2) & 1)
    return (bslot * 8 + ((r >> 1) & 7)) * 2 + (r & 1)


def _rfast80_chunk(idx):
    """Inverse of _rfast80_at for the loads: chunk id -> (k, row pair p)."""
    bs = idx >> 3
    kr = (bs >> 1) // 5
    b = (bs >> 1) - 5 * kr
    return (kr & 3) + 8 * (kr >> 2) + 4 * (bs & 1), 8 * b + (idx & 7)


def _direct_stage_map(R, rfast):
    """LDS slot (in doubles) of each tile element (r, k), r < R, k < 16, as the
    direct-load plan of gemm_direct.hip places it: thread t, load q writes 16
    bytes at byte 16 (512 q + t); kfast slot = 16 r + 2 (c ^ f(r)) + k % 2,
    rfast slot = R k + 2 (p ^ sw(k)) + r % 2 (R = 80: _rfast80_at).  Loads past
    the tile's 8 R chunks (R = 80: 1024 slots for 640 chunks) fill padding and
    hold no element."""
    slot = {}
    nch = 8 * R
    for q in range(-(-nch // 512)):
        for t in range(512):
            idx = q * 512 + t
            if idx >= nch:
                continue
            if rfast and R == 80:
                k, p = _rfast80_chunk(idx)
                for e in range(2):
                    slot[(2 * p + e, k)] = 2 * idx + e
            elif rfast:
                k, pc = divmod(idx, R // 2)
                p = pc ^ _rfast_swz(R, k)
                assert p < R // 2
                for e in range(2):
                    slot[(2 * p + e, k)] = 2 * idx + e
            else:
                rl, pc = idx >> 3, idx & 7
                c = pc ^ ((rl ^ (rl >> 3)) & 7)
                for e in range(2):
                    slot[(rl, 2 * c + e)] = 2 * idx + e
    return slot


@pytest.mark.parametrize('R', [64, 80, 128])
@pytest.mark.parametrize('rfast', [0, 1])
def test_direct_engine_lds_layout_round_trip(R, rfast):
    """The direct-load plan stores every (row, k) of a chunk exactly once, and
    the fragment reads of the kernel (lane l: row 16 t + l % 16, k = 4 (l / 16)
    + s) find each element where it was stored."""
    slot = _direct_stage_map(R, rfast)
    assert len(slot) == R * 16 and sorted(slot.values()) == list(range(R * 16))
    for r in range(R):
        for k in range(16):
            g4, s = (k // 4) * 4, k % 4
            if rfast and R == 80:
                read = _rfast80_at(k, r)
            elif rfast:
                c = (r >> 1) ^ _rfast_swz(R, k)
                read = k * R + 2 * c + (r & 1)
            else:
                f = (r ^ (r >> 3)) & 7
                chunk = ((g4 >> 1) + (s >> 1)) ^ f
                read = r * 16 + 2 * chunk + (s & 1)
            assert read == slot[(r, k)], (r, k)


def test_direct_engine_rfast80_groups_in_distinct_half_windows():
    """80-row row-contiguous FP64 stage: the lane groups 0/1 (and 2/3) of every
    fragment read (k = 4 g + s, rows r16 .. r16 + 15) read 128 contiguous bytes
    each, in different 128-byte halves of the 256-byte bank window, so one half
    wave's 8-byte reads meet no bank twice."""
    for r16 in range(0, 80, 16):
        for s in range(4):
            for g in (0, 2):
                for gg in (g, g + 1):
                    offs = sorted(_rfast80_at(4 * gg + s, r16 + i) * 8 for i in range(16))
                    assert offs == list(range(offs[0], offs[0] + 128, 8))
                h0 = (_rfast80_at(4 * g + s, r16) * 8 // 128) % 2
                h1 = (_rfast80_at(4 * (g + 1) + s, r16) * 8 // 128) % 2
                assert h0 != h1, (r16, s, g)


@pytest.mark.parametrize('o,v,i,j,kp,k', [(5, 12, 0, 1, 2, 3), (5, 12, 1, 1, 2, 4), (4, 10, 0, 2, 2, 3),
                                          (4, 9, 0, 0, 1, 3)])
def test_paired_adjoint_descriptors_match_the_unpaired_ones(o, v, i, j, kp, k):
    """The two-stream walk's paired launches for consecutive triples (i, j, kp),
    (i, j, k) (the second triple's ovvv and hole-term t2 adjoints into index i
    and j as four-segment launches over both triples' Wbar) leave Vpb and T2mb
    as the two triples' unpaired launches do, with every address in bounds and
    no write race inside a launch."""
    p = make_problem(o, v, 3 * v, seed=o * 31 + v + k)
    rng = np.random.default_rng(o + v + kp)
    w1, w2 = rng.normal(size=v ** 3), rng.normal(size=v ** 3)
    sz = sizes(o, v)
    sz['W2'] = v ** 3
    mem = Memory(names=NAMES + ('W2',), array_sizes=sz)
    _fill(mem, p)
    mem.arrays['W'][:] = w1
    mem.arrays['W2'][:] = w2
    for d in record_paired(o, v, i, j, kp, k):
        mem.run(d)
    ref = Memory(o, v)
    _fill(ref, p)
    ref.arrays['W'][:] = w1
    for d in record(o, v, i, j, kp, with_vjp=True)[3:]:
        ref.run(d)
    ref.arrays['W'][:] = w2
    for d in record(o, v, i, j, k, with_vjp=True)[3:]:
        ref.run(d)
    for name in ('Vpb', 'T2mb'):
        a, b = mem.arrays[name], ref.arrays[name]
        assert np.abs(b).max() > 0
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), name
