"""pamd_ccsdt_energy_vjp_noalloc (library-owned scratch, no device allocation)
against pamd_ccsdt_energy_vjp with a caller workspace, through the C ABI: two
calls in a row, the second on a smaller problem, each within 1e-10 Eh in E(T)
and 1e-9 relative per cotangent block (ovvv_bar and t2_bar in the split form:
1e-6, since the workspace entry point's two-stream walk pairs the ovvv and
hole-term t2 adjoints of consecutive triples into other FP32 sums); and a
problem larger than the scratch is rejected with PAMD_ERR_WORKSPACE."""
import ctypes

import numpy as np
import pytest
import torch

from pyscfad_amd import _lib
from pyscfad_amd.cc import _ops
from pyscfad_amd.gen import make_problem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def _noalloc(inp):
    lib = _lib.load()
    out = _ops._cotangent_outputs(inp)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    st = lib.pamd_ccsdt_energy_vjp_noalloc(ctypes.byref(inp.cfg), *inp.args(),
                                           ctypes.c_double(0.9), _ops._ptr(e),
                                           *_ops._cotangent_ptrs(out), _ops._stream(inp.device))
    torch.cuda.synchronize()
    return st, e, out


@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
def test_noalloc_matches_workspace_entry_point(precision, cuda):
    for o, v, seed in ((8, 40, 3), (5, 24, 4)):
        print(f'noalloc {precision} o{o} v{v}', flush=True)
        p = make_problem(o, v, 4 * v, seed=seed, fvo_scale=1e-3)
        inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, p.ovvv_packed(), p.ovoo, p.ovov, p.fvo,
                          precision=precision, device=cuda)
        e_ref, g_ref = _ops.energy_vjp(inp, et_bar=0.9)
        st, e, g = _noalloc(inp)
        assert st == _lib.PAMD_OK
        assert abs(float(e.item()) - float(e_ref.item())) <= 1e-10
        for k, ref in g_ref.items():
            tol = 1e-6 if precision != 'fp64' and k in ('ovvv_bar', 't2_bar') else 1e-9
            assert _rel(g[k], ref) <= tol, (k, _rel(g[k], ref))


def test_noalloc_rejects_problems_beyond_the_scratch(cuda):
    lib = _lib.load()
    cap = lib.pamd_ccsdt_noalloc_scratch_bytes()
    cfg = _lib.make_config(16, 64, ovvv_packed=True)
    assert _lib.workspace_size(cfg, True) > cap
    p = make_problem(16, 64, 256, seed=1)
    inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, p.ovvv_packed(), p.ovoo, p.ovov, None, device=cuda)
    st, _e, _g = _noalloc(inp)
    assert st == _lib.PAMD_ERR_WORKSPACE
