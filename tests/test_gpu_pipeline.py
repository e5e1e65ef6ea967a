"""The adjoint walk with two triples in flight (pamd_ccsdt_energy_vjp) against
the walk with one triple in flight (pamd_ccsdt_energy_vjp_noalloc, which never
pipelines): every accumulator shared by the triples is written by one stream in
walk order, so the energies and cotangent blocks are bitwise equal, except that
in the split forms the two-stream walk also pairs the ovvv and hole-term t2
adjoints of consecutive triples (i, j, k), (i, j, k') into one launch each for
Vpb[i], Vpb[j], T2mb[i], T2mb[j]: ovvv_bar and t2_bar then hold the same terms
in other FP32 sums (within 1e-6 relative, about 1e-8 measured); FP64 never
pairs."""
import ctypes

import pytest
import torch

from pyscfad_amd import _lib
from pyscfad_amd.cc import _ops
from pyscfad_amd.gen import make_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
@pytest.mark.parametrize('o,v,seed', [(8, 40, 5), (5, 24, 6)])
def test_pipelined_walk_matches_the_one_stream_walk(o, v, seed, precision, cuda):
    print(f'pipeline vs one stream: {precision} o{o} v{v}', flush=True)
    p = make_problem(o, v, 4 * v, seed=seed, fvo_scale=1e-3)
    inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, p.ovvv_packed(), p.ovoo, p.ovov, p.fvo,
                      precision=precision, device=cuda)
    e_ref, g_ref = _ops.energy_vjp(inp, et_bar=0.7)
    lib = _lib.load()
    out = _ops._cotangent_outputs(inp)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    st = lib.pamd_ccsdt_energy_vjp_noalloc(ctypes.byref(inp.cfg), *inp.args(),
                                           ctypes.c_double(0.7), _ops._ptr(e),
                                           *_ops._cotangent_ptrs(out), _ops._stream(inp.device))
    torch.cuda.synchronize()
    assert st == _lib.PAMD_OK
    assert torch.equal(e, e_ref.reshape(e.shape))
    for k, ref in g_ref.items():
        if precision != 'fp64' and k in ('ovvv_bar', 't2_bar'):
            rel = float((out[k] - ref).norm() / ref.norm())
            assert rel <= 1e-6, (k, rel)
        else:
            assert torch.equal(out[k], ref), k
