"""Split-BF16x3 against plain FP64 E(T) on one round-robin shard of the triples.

    python -m pyscfad_amd.cc.precision_check [--nocc 60 --nvir 600 --naux 2400]
        [--nshards 16 --shard 0] [--out report.json]

At sizes where a whole FP64 E(T) does not fit one run (nocc=60/nvir=600 takes
about 35 minutes in FP64), the error of the split-BF16x3 form is measured on a
shard of the occupied triples: every ``nshards``-th triple of the walk, the
per-GPU share of a sharded run.  Both precisions evaluate the same triples of
the same seeded problem.  The report holds the shard energies, their
difference, the time per triple of each form and the whole-E(T) times they
extrapolate to (x nshards).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import torch

from .. import _lib
from ..gen import make_problem
from . import _ops


def run(nocc, nvir, naux, nshards, shard, seed=0, log=print):
    dev = torch.device('cuda', 0)
    t0 = time.perf_counter()
    p = make_problem(nocc, nvir, naux, seed=seed, device=dev)
    idx, idy = torch.tril_indices(nvir, nvir, device=dev)
    ovvv = p.ovvv[:, :, idx, idy].contiguous()
    del p.ovvv
    torch.cuda.synchronize()
    log(f'problem o{nocc}/v{nvir}/naux{naux} built in {time.perf_counter() - t0:.1f} s')
    ntrip = _ntrip(nocc)
    n_shard = len(range(shard, ntrip, nshards))
    rep = {'workload': f'E(T) nocc={nocc} nvir={nvir} naux={naux} seed={seed}',
           'nshards': nshards, 'shard': shard, 'triples_total': ntrip, 'triples_in_shard': n_shard}
    _lib.profile_progress(max(1, n_shard // 8))
    for prec in ('bf16x3', 'fp64'):
        inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, ovvv, p.ovoo, p.ovov, None, precision=prec,
                          device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = float(_ops.energy_shard(inp, shard, nshards).item())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rep[prec] = {'e_shard': e, 'sec': dt, 'ms_per_triple': 1e3 * dt / n_shard,
                     'sec_whole_e_t_extrapolated': dt * ntrip / n_shard}
        log(f'{prec}: shard E = {e:.15e} in {dt:.1f} s')
        del inp
        torch.cuda.empty_cache()
    _lib.profile_progress(0)
    d = abs(rep['bf16x3']['e_shard'] - rep['fp64']['e_shard'])
    rep['abs_dE_shard'] = d
    rep['rel_dE_shard'] = d / abs(rep['fp64']['e_shard'])
    rep['speedup'] = rep['fp64']['sec'] / rep['bf16x3']['sec']
    log(json.dumps(rep))
    return rep


def _ntrip(nocc):
    import ctypes
    n = ctypes.c_longlong(0)
    _lib.check(_lib.load().pamd_ccsdt_num_triples(nocc, ctypes.byref(n)), 'pamd_ccsdt_num_triples')
    return n.value


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--nocc', type=int, default=60)
    ap.add_argument('--nvir', type=int, default=600)
    ap.add_argument('--naux', type=int, default=2400)
    ap.add_argument('--nshards', type=int, default=16)
    ap.add_argument('--shard', type=int, default=0)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--out', default=None)
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print('precision_check: no GPU visible', file=sys.stderr)
        return 1
    rep = run(args.nocc, args.nvir, args.naux, args.nshards, args.shard, args.seed,
              log=lambda s: print(s, flush=True))
    if args.out:
        with open(args.out, 'w') as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == '__main__':
    sys.exit(main())
