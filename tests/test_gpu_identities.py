"""Exact identities of E(T) and its adjoint on the GPU, at sizes beyond the
oracle's reach (the oracle holds o^3 v^3 arrays).  They hold for any inputs:

* D depends on orbital energies only through e_i + e_j + e_k - e_a - e_b - e_c:
  sum(dE/de) = 0 (shift invariance) and <e, dE/de> = -E (degree -1);
* with fvo = 0, E = A + B where A = E(t1 = 0) is quadratic in W and B is linear
  in t1 and in ovov, W being linear in t2 and in (ovvv, ovoo):
  <t1, dE/dt1> = <ovov, dE/dovov> = B and
  <t2, dE/dt2> = <ovvv, dE/dovvv> + <ovoo, dE/dovoo> = 2A + B.

A dropped term, a wrong sign or a transposed operand in any adjoint block
breaks one of them.
"""
import pytest
import torch

from pyscfad_amd.cc import _ops
from pyscfad_amd.gen import make_problem

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return float((a * b).sum())


@pytest.mark.parametrize('o,v,precision,tol', [(16, 160, 'fp64', 1e-11), (16, 160, 'bf16x3', 1e-6),
                                               (24, 240, 'fp64', 1e-11),
                                               (24, 240, 'bf16x3', 1e-6)])
def test_exact_identities_mid_size(o, v, precision, tol, cuda):
    print(f'identities o{o} v{v} {precision}', flush=True)
    p = make_problem(o, v, 4 * v, seed=21, device=cuda)
    idx, idy = torch.tril_indices(v, v, device=cuda)
    ovvv = p.ovvv[:, :, idx, idy].contiguous()
    del p.ovvv
    inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, ovvv, p.ovoo, p.ovov, None, precision=precision,
                      device=cuda)
    e_t, g = _ops.energy_vjp(inp, et_bar=1.0)
    e = float(e_t.item())
    inp0 = _ops.Inputs(p.mo_energy, torch.zeros_like(p.t1), p.t2, ovvv, p.ovoo, p.ovov, None,
                       precision=precision, device=cuda)
    A = float(_ops.energy(inp0).item())
    B = e - A
    mo = p.mo_energy
    scale = abs(e)
    assert abs(float(g['mo_energy_bar'].sum())) <= tol * scale / float(mo.abs().mean())
    assert abs(_dot(mo, g['mo_energy_bar']) + e) <= tol * scale
    assert abs(_dot(p.t1, g['t1_bar']) - B) <= tol * scale
    assert abs(_dot(p.ovov, g['ovov_bar']) - B) <= tol * scale
    assert abs(_dot(p.t2, g['t2_bar']) - (2 * A + B)) <= tol * scale
    assert abs(_dot(ovvv, g['ovvv_bar']) + _dot(p.ovoo, g['ovoo_bar']) - (2 * A + B)) <= tol * scale


@pytest.mark.parametrize('o,v', [(16, 160), (24, 240)])
def test_bf16x3_blocks_against_fp64(o, v, cuda):
    """Split-BF16x3 E(T)+VJP against the FP64 kernel on the same inputs, block
    by block: |dE| < 1e-7 Eh and every cotangent block within 1e-6 relative
    (2-norm).  At v = 160 and 240 the virtual-side contractions run on the
    128x80 / 80x128 two-workgroup tiles; at o = 24 the occupied-side ones on the
    128x48 tile, at o = 16 on the one-workgroup 256x64 fallback."""
    print(f'bf16x3 vs fp64 blocks o{o} v{v}', flush=True)
    p = make_problem(o, v, 4 * v, seed=5, fvo_scale=1e-3, device=cuda)
    idx, idy = torch.tril_indices(v, v, device=cuda)
    ovvv = p.ovvv[:, :, idx, idy].contiguous()
    del p.ovvv
    fvo = p.fock[o:, :o].contiguous()
    out = {}
    for prec in ('fp64', 'bf16x3'):
        inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, ovvv, p.ovoo, p.ovov, fvo, precision=prec,
                          device=cuda)
        out[prec] = _ops.energy_vjp(inp, et_bar=1.0)
        del inp
    e64, g64 = out['fp64']
    e16, g16 = out['bf16x3']
    assert abs(float(e16.item()) - float(e64.item())) < 1e-7
    for key, ref in g64.items():
        err = float((g16[key] - ref).norm() / ref.norm())
        assert err < 1e-6, (key, err)
