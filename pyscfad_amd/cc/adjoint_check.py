"""Adjoint check of the (T) kernels by central finite differences.

    python -m pyscfad_amd.cc.adjoint_check [--nocc 24 --nvir 240 --naux 960]
        [--directions 64] [--h 1e-5] [--precision fp64 bf16x3] [--nshards 1]
        [--out report.json]

For each seeded unit direction d in the joint input space (mo_energy, t1, t2,
packed ovvv, ovoo, ovov, fvo), compares the directional derivative from one
VJP of each precision, <grad E, d>, with the central difference
(E(x + h d) - E(x - h d)) / (2 h) of the plain FP64 energy.  (The split-BF16x3
energy is exact to ~1e-10 relative but not smooth at that level, so its own
difference quotient would measure rounding, not the derivative.)  Also
reports the peak device memory of each VJP call against what storing W and V
of every occupied triple would take.

Error measures per precision, over the directions (fd = difference quotient,
ana = <g, d> of that precision's VJP, rms = root mean square of the FP64 VJP's
directional derivatives over the sampled directions):
  rel_err        |fd - ana| / |fd|, relative to each direction's own derivative
                 (large where that derivative happens to be small);
  err_vs_rms     |fd - ana| / rms;
  vs_fp64_rms    (precisions other than fp64) |ana - ana_fp64| / rms: the error
                 against the FP64 VJP, free of the difference quotient's own
                 truncation and rounding error.
The per-direction values are kept in the report.

With ``nshards`` > 1 the energy and its VJP are those of round-robin shard 0
of the occupied triples (pamd_ccsdt_energy_shard / _energy_vjp_shard): a
function of the same inputs with its own exact gradient, so the same check
runs at full input size for 1/nshards of the cost.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from ..gen import make_problem
from . import _ops

BLOCKS = ('mo_energy', 't1', 't2', 'ovvv', 'ovoo', 'ovov', 'fvo')


def _inputs(x, precision, dev):
    return _ops.Inputs(x['mo_energy'], x['t1'], x['t2'], x['ovvv'], x['ovoo'], x['ovov'],
                       x['fvo'], precision=precision, device=dev)


def _energy(x, nshards, dev):
    inp = _inputs(x, 'fp64', dev)
    if nshards > 1:
        return float(_ops.energy_shard(inp, 0, nshards).item())
    return float(_ops.energy(inp).item())


def run(nocc, nvir, naux, directions, h, precisions, seed=0, dirseed=1234, log=print,
        nshards=1):
    dev = torch.device('cuda', 0)
    p = make_problem(nocc, nvir, naux, seed=seed, fvo_scale=1e-3, device=dev)
    iu, ju = torch.tril_indices(nvir, nvir, device=dev)
    x = dict(mo_energy=p.mo_energy, t1=p.t1, t2=p.t2, ovvv=p.ovvv[:, :, iu, ju].contiguous(),
             ovoo=p.ovoo, ovov=p.ovov, fvo=p.fock[nocc:, :nocc].contiguous())
    del p
    gen = torch.Generator(device=dev).manual_seed(dirseed)
    report = {'workload': f'o{nocc}/v{nvir}/naux{naux}', 'h': h, 'directions': directions,
              'fd': 'central difference of the FP64 energy', 'nshards': nshards,
              'precisions': {}}
    grads, meta = {}, {}
    o, v = nocc, nvir
    ntrip = o * (o + 1) * (o + 2) // 6 - o
    for prec in precisions:
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        inp = _inputs(x, prec, dev)
        t0 = time.perf_counter()
        if nshards > 1:
            e0, g = _ops.energy_vjp_shard(inp, 0, nshards, et_bar=1.0)
        else:
            e0, g = _ops.energy_vjp(inp, et_bar=1.0)
        torch.cuda.synchronize()
        meta[prec] = {'e_t': float(e0.item()), 'vjp_s': time.perf_counter() - t0,
                      'peak_device_bytes_vjp': int(torch.cuda.max_memory_allocated(dev) - base),
                      'store_everything_bytes': int(2 * 8 * ntrip * v ** 3)}
        grads[prec] = {'mo_energy': g['mo_energy_bar'], 't1': g['t1_bar'], 't2': g['t2_bar'],
                       'ovvv': g['ovvv_bar'], 'ovoo': g['ovoo_bar'], 'ovov': g['ovov_bar'],
                       'fvo': g['fvo_bar']}
        del inp, g
    rel = {prec: [] for prec in precisions}
    vals = {'fd': [], **{prec: [] for prec in precisions}}
    t0 = time.perf_counter()
    for q in range(directions):
        d = {kk: torch.randn(x[kk].shape, generator=gen, dtype=torch.float64, device=dev)
             for kk in BLOCKS}
        nrm = torch.sqrt(sum((t * t).sum() for t in d.values()))
        d = {kk: t / nrm for kk, t in d.items()}
        ep = _energy({kk: x[kk] + h * d[kk] for kk in BLOCKS}, nshards, dev)
        em = _energy({kk: x[kk] - h * d[kk] for kk in BLOCKS}, nshards, dev)
        fd = (ep - em) / (2 * h)
        vals['fd'].append(fd)
        msg = []
        ana = {prec: float(sum((grads[prec][kk] * d[kk]).sum() for kk in BLOCKS))
               for prec in precisions}
        for prec in precisions:
            vals[prec].append(ana[prec])
            rel[prec].append(abs(fd - ana[prec]) / max(abs(fd), 1e-300))
            msg.append(f'{prec} {ana[prec]:.10e} (rel {rel[prec][-1]:.2e})')
        if q % 8 == 0 or q == directions - 1:
            log(f'direction {q + 1}/{directions}: fd {fd:.10e}; ' + '; '.join(msg))
    t_fd = time.perf_counter() - t0
    summ = summarise(vals, precisions)
    for prec in precisions:
        m = meta[prec]
        m.update({'rel_err_max': max(rel[prec]), 'rel_err_median': float(np.median(rel[prec])),
                  'fd_s': t_fd,
                  'memory_ratio': m['store_everything_bytes'] / max(m['peak_device_bytes_vjp'], 1)})
        m.update(summ[prec])
        report['precisions'][prec] = m
        log(json.dumps({prec: m}))
    report['directional_derivatives'] = vals
    return report


def summarise(vals, precisions):
    """Error measures from the per-direction values {'fd': [...], prec: [...]}:
    err_vs_rms and, for precisions other than fp64 when fp64 was run,
    vs_fp64_rms (max and median of each; definitions in the module docstring)."""
    fd = np.asarray(vals['fd'], dtype=float)
    ref = np.asarray(vals['fp64' if 'fp64' in vals else precisions[0]], dtype=float)
    rms = float(np.sqrt(np.mean(ref * ref)))
    out = {}
    for prec in precisions:
        a = np.asarray(vals[prec], dtype=float)
        e = np.abs(fd - a) / rms
        m = {'rms_derivative': rms, 'err_vs_rms_max': float(e.max()),
             'err_vs_rms_median': float(np.median(e))}
        if prec != 'fp64' and 'fp64' in vals:
            e64 = np.abs(a - ref) / rms
            m.update({'vs_fp64_rms_max': float(e64.max()), 'vs_fp64_rms_median': float(np.median(e64))})
        out[prec] = m
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--nocc', type=int, default=24)
    ap.add_argument('--nvir', type=int, default=240)
    ap.add_argument('--naux', type=int, default=960)
    ap.add_argument('--directions', type=int, default=64)
    ap.add_argument('--h', type=float, default=1e-5)
    ap.add_argument('--precision', nargs='+', default=['fp64', 'bf16x3'])
    ap.add_argument('--nshards', type=int, default=1)
    ap.add_argument('--out', default=None)
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print('adjoint_check: no GPU visible', file=sys.stderr)
        return 1
    rep = run(args.nocc, args.nvir, args.naux, args.directions, args.h, args.precision,
              log=lambda s: print(s, flush=True), nshards=args.nshards)
    if args.out:
        with open(args.out, 'w') as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == '__main__':
    sys.exit(main())
