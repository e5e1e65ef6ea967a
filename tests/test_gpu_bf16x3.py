"""Split-BF16x3 (T): accuracy against the FP64 path and the CPU oracle.

The split form runs the contractions as bf16 MFMA products of (hi, mid, lo)
pieces with FP32 tile accumulation; the epilogue, denominators and reductions
stay FP64.  Bars (relative to the FP64 result): E(T) within 1e-6 relative, each
gradient block within 1e-5 relative (Frobenius).  At nocc=40/nvir=400 the
target is |dE| < 1e-7 Eh and 1e-6 relative per block (bench.py --compare).
"""
import numpy as np
import pytest

from pyscfad_amd.gen import make_problem
from pyscfad_amd.cc import _ops
from oracle import ccsd_t_oracle as O

pytestmark = pytest.mark.gpu

BLOCKS = ('t1_bar', 't2_bar', 'fvo_bar', 'mo_energy_bar', 'ovoo_bar', 'ovov_bar', 'ovvv_bar')


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize('o,v,fvo,packed', [(4, 16, 0.0, True), (6, 24, 1e-3, False),
                                             (8, 40, 0.0, True), (4, 48, 1e-3, True)])
def test_bf16x3_vs_oracle(o, v, fvo, packed, cuda):
    p = make_problem(o, v, 4 * v, seed=17, fvo_scale=fvo)
    print(f'case o{o}v{v} fvo={fvo} packed={packed}', flush=True)
    args = (p.mo_energy, p.t1, p.t2, p.ovvv, p.ovoo, p.ovov, p.fvo)
    e_ref, g_ref = O.energy_vjp(*args, et_bar=1.0)
    ovvv = p.ovvv_packed() if packed else p.ovvv
    inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, ovvv, p.ovoo, p.ovov,
                      p.fvo if fvo else None, precision='bf16x3', device=cuda)
    e = float(_ops.energy(inp).item())
    e2, g = _ops.energy_vjp(inp, et_bar=1.0)
    print(f'  E rel err {abs(e - e_ref) / abs(e_ref):.3e}', flush=True)
    assert abs(e - e_ref) <= 1e-6 * abs(e_ref)
    assert abs(float(e2.item()) - e) <= 1e-12 * abs(e_ref)
    for k in BLOCKS:
        ref = O.fold_ovvv_packed(g_ref[k]) if (k == 'ovvv_bar' and packed) else g_ref[k]
        err = _rel(g[k].cpu().numpy(), ref)
        print(f'  {k} rel err {err:.3e}', flush=True)
        assert err <= 1e-5, k


# v = 128: several full 256x64 / 64x256 tiles; v = 80: the 16x16-MFMA tiles
# (256x80, 80x256, 256x48) that the engine picks for v = 400
@pytest.mark.parametrize('o,v', [(12, 128), (6, 80)])
def test_bf16x3_vs_fp64_larger(o, v, cuda):
    p = make_problem(o, v, 4 * v, seed=3, device=cuda)
    ins = {}
    for prec in ('fp64', 'bf16x3'):
        ins[prec] = _ops.Inputs(p.mo_energy, p.t1, p.t2, p.ovvv, p.ovoo, p.ovov, None,
                                precision=prec, device=cuda)
    e64, g64 = _ops.energy_vjp(ins['fp64'])
    e16, g16 = _ops.energy_vjp(ins['bf16x3'])
    de = abs(float(e16.item()) - float(e64.item()))
    print(f'o{o}v{v} E64={float(e64.item()):.12e} |dE|={de:.3e}', flush=True)
    assert de <= 1e-6 * abs(float(e64.item()))
    for k in BLOCKS:
        err = _rel(g16[k].cpu().numpy(), g64[k].cpu().numpy())
        print(f'  {k} rel err {err:.3e}', flush=True)
        assert err <= 1e-5, k
