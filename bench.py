#!/usr/bin/env python
"""Benchmark of the closed-shell (T) energy + VJP on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision bf16x3|fp64]
                    [--compare shard|full|none] [--budget SECONDS]

One step = one E(T)+VJP pass (E and every cotangent block, et_bar = 1) over the
seeded synthetic problem nocc=40, nvir=400, naux=1600 (packed ovvv, the DF layout),
with all inputs already resident in HBM when the timer starts.  The timed form
is the split-BF16x3 one by default.

The run keeps to a wall-clock budget (``--budget`` seconds from the start of
the process, default 480): the W warmup passes run on a small problem (the HIP
code needs no compilation, so they only load the library and touch every
kernel), then up to K full steps are timed -- as many as fit the budget, at
least one -- and the line reports the number actually timed.  The FP64
comparison runs the plain FP64 form and the timed form on the same
round-robin shard of the occupied triples (``--compare-shards``, default 1/64 of
the walk, through pamd_ccsdt_energy_vjp_shard): |dE| and the relative error of
every gradient block on that shard, the time of both forms on it, and the FP64
whole-step time it extrapolates to.  ``--compare full`` runs a whole FP64 step
instead (about 400 s at the default size).

Prints one JSON line: seconds per E(T)+VJP, the roofline of the dominant kernel
(the contraction engine; its launches are timed with hipEvents on their own
stream for one occupied triple in --profile-every), and the reference's own C
(oracle/_ref, built from the reference sources) timed on this host's cores on a
bounded sample and scaled by algorithmic flops.  Progress lines go to stderr.
"""
from __future__ import annotations

import time

_T_START = time.time()   # the wall-clock budget counts from here

import argparse  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402
import platform  # noqa: E402
import sys  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {'fp64': 78.6, 'bf16': 2500.0, 'fp32': 157.3}  # dense MFMA TFLOP/s, MI355X


def alg_flops(o, v):
    """E + VJP algorithmic flops: 3 x 2 o^3 v^3 (o + v)."""
    return 3.0 * 2.0 * o ** 3 * v ** 3 * (o + v)


def cpu_model():
    """Host CPU model name and the number of logical CPUs this process may use."""
    name = platform.processor() or ''
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    name = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return name, ncpu


def cpu_baseline(o_full, v_full, sample, threads):
    """Reference C (oracle/_ref) E+VJP on a smaller seeded problem, scaled by
    algorithmic flops.  The flat-GFLOP/s premise of that scaling is measured
    over o8/v80 .. o16/v160 in profiles/cpu_baseline_scaling.md."""
    from oracle import ref_harness as R
    from pyscfad_amd.gen import make_problem
    so, sv = sample
    if not R.available():
        return None
    os.environ.setdefault('OMP_NUM_THREADS', str(threads))
    p = make_problem(so, sv, 4 * sv, seed=0)
    args = (p.mo_energy, p.t1, p.t2, p.ovvv, p.ovoo, p.ovov, p.fvo)
    t0 = time.perf_counter()
    R.energy(*args)
    R.energy_vjp(*args)
    dt = time.perf_counter() - t0
    scale = alg_flops(o_full, v_full) / alg_flops(so, sv)
    model, ncpu = cpu_model()
    return {'value': dt * scale, 'unit': 's', 'cores': threads, 'kind': 'reference',
            'sample': (f'reference C E(T)+VJP (lnoccsdt_contract + ccsd_t_energy_vjp, one '
                       f'block) at o{so}/v{sv}/naux{4 * sv} took {dt:.2f} s on {threads} '
                       f'OpenMP threads (MKL sequential); scaled by algorithmic flops '
                       f'x{scale:.4g} to o{o_full}/v{v_full}'),
            'sample_sec': dt, 'gflops': alg_flops(so, sv) / dt / 1e9,
            'host_cpu': model, 'host_logical_cpus': ncpu,
            'scaling_check': 'profiles/cpu_baseline_scaling.md'}


def engine_roofline(prof, precision):
    """Roofline of the contraction engine from the sampled hipEvent timings."""
    gemm = [x for k, x in prof.items() if k.startswith('gemm_') and x['launches']]
    g_ms = sum(x['ms'] for x in gemm)
    g_fl = sum(x['flops'] for x in gemm)   # FP64-equivalent: 2 M N K per launch
    g_n = sum(x['launches'] for x in gemm)
    fp64_equiv = g_fl / (g_ms * 1e-3) / 1e12 if g_ms > 0 else None
    if precision == 'bf16x3':
        # six bf16 MFMA partial products per FP64-equivalent multiply-add
        kernel, mfma_factor, peak = 'k_gemm_b3p (k_gemm_b3n for operands without 16-byte runs)', 6.0, PEAKS['bf16']
    elif precision == 'fp32x2':
        # three fp32 MFMA products (hi*hi, hi*lo, lo*hi) per multiply-add
        kernel, mfma_factor, peak = 'k_gemm_f64 (FP32 hi/lo mode)', 3.0, PEAKS['fp32']
    else:
        kernel, mfma_factor, peak = 'k_gemm_f64', 1.0, PEAKS['fp64']
    achieved = fp64_equiv * mfma_factor if fp64_equiv else None
    traffic, traffic_src = engine_traffic(precision)
    roofline = {'bound': 'mfma', 'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s',
                'frac': (achieved / peak) if achieved else None, 'traffic': traffic,
                'traffic_unit': 'bytes per launch' if traffic else None,
                'traffic_source': traffic_src,
                'kernel': kernel, 'mfma_flops_per_alg_flop': mfma_factor,
                'fp64_equiv_tflops': fp64_equiv,
                'frac_fp64_peak': (fp64_equiv / PEAKS['fp64']) if fp64_equiv else None,
                'launches_timed': g_n,
                'avg_launch_ms': g_ms / g_n if g_n else None,
                'flops_per_launch': g_fl * mfma_factor / g_n if g_n else None}
    return roofline


# rocprofv3 FETCH_SIZE / WRITE_SIZE summaries of the engine launches: the
# split-BF16x3 one at the default nocc=40 (its tile shapes for the occupied-side
# contractions depend on nocc), the FP64 one at nocc=10 (the FP64 engine runs the
# same two tile shapes at every size)
TRAFFIC_FILES = {p: os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', f)
                 for p, f in (('bf16x3', 'traffic_e_t_vjp_o40v400_bf16x3.json'),
                              ('fp64', 'traffic_e_t_vjp_o10v400_fp64.json'))}


def engine_traffic(precision):
    """HBM bytes per launch of the contraction engine (dispatch-weighted mean
    over its tile shapes) from the committed rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE summary, or None when that file is absent."""
    path = TRAFFIC_FILES.get(precision)
    if path is None:
        return None, None
    prefix = 'k_gemm_b3' if precision == 'bf16x3' else 'k_gemm_f64'
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    ks = [x for k, x in t['kernels'].items() if k.startswith(prefix)]
    n = sum(x['dispatches'] for x in ks)
    if not n:
        return None, None
    b = sum(x['dispatches'] * (x['hbm_read_bytes'] + x['hbm_write_bytes']) for x in ks) / n
    return b, f"rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, {t['workload']} (profiles/)"


def df_measure(o, v, naux, seed, dev, inp, grads):
    """Time pamd_df_eris / pamd_df_eris_vjp on the workload's own DF factors: the
    forward rebuilds the (T) blocks, the adjoint folds the step's cotangents
    into dE/dLpq.  Reported beside the (T) step, not part of it."""
    import torch
    from pyscfad_amd.cc import df as dfm
    from pyscfad_amd.gen import make_L
    L = make_L(o, v, naux, seed=seed, device=dev)
    n = o + v
    npair = v * (v + 1) // 2
    f_fwd = 2.0 * naux * o * v * (npair + o * o + o * v)
    f_vjp = 2.0 * naux * o * v * (o * o + (npair + o * o + 2 * o * v) + npair)
    dfm.df_eris(L[:, :min(n, 8), :min(n, 8)].contiguous(), min(o, 2))  # load, warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = dfm.df_eris(L, o)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    diff = max(float((out[k] - ref).abs().max() / ref.abs().max())
               for k, ref in (('ovvv', inp.ovvv), ('ovoo', inp.ovoo), ('ovov', inp.ovov)))
    del out
    t0 = time.perf_counter()
    Lb = dfm.df_eris_vjp(L, o, grads['ovvv_bar'], grads['ovoo_bar'], grads['ovov_bar'])
    torch.cuda.synchronize()
    t_vjp = time.perf_counter() - t0
    del Lb, L
    return {'workload': f'Lpq ({naux}, {n}, {n}) -> packed ovvv, ovoo, ovov and back',
            'fwd_s': t_fwd, 'vjp_s': t_vjp, 'fwd_tflops': f_fwd / t_fwd / 1e12,
            'vjp_tflops': f_vjp / t_vjp / 1e12, 'frac_fp64_peak_fwd': f_fwd / t_fwd / 1e12 / PEAKS['fp64'],
            'max_rel_diff_vs_generator': diff}


def _log(msg):
    print(f'bench [{time.time() - _T_START:6.1f} s] {msg}', file=sys.stderr, flush=True)


def grad_errors(g, g64):
    import torch
    return {k: float(torch.linalg.norm(g[k] - g64[k]) / torch.linalg.norm(g64[k])) for k in g64}


def compare_on_shard(inp, ref_inp, nshards, prof_every, precision):
    """The timed form and plain FP64 on the same round-robin shard of the
    occupied triples (pamd_ccsdt_energy_vjp_shard)."""
    import torch
    from pyscfad_amd import _lib
    from pyscfad_amd.cc import _ops
    n_trip = (inp.nocc * (inp.nocc + 1) * (inp.nocc + 2)) // 6 - inp.nocc
    n_shard = len(range(0, n_trip, nshards))
    out = {'mode': 'shard', 'nshards': nshards, 'shard': 0, 'triples_in_shard': n_shard,
           'triples_total': n_trip}
    res = {}
    for tag, x in ((precision, inp), ('fp64', ref_inp)):
        ws = x.workspace(True)
        _lib.profile_reset()
        _lib.profile_enable(prof_every)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e, g = _ops.energy_vjp_shard(x, 0, nshards, et_bar=1.0, workspace=ws)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _lib.profile_enable(0)
        res[tag] = (float(e.item()), g, dt, engine_roofline(_lib.profile_read(), tag))
        del ws
        torch.cuda.empty_cache()
        _log(f'compare: {tag} shard 0/{nshards} ({n_shard} triples) E+VJP {dt:.2f} s')
    e_t, g_t, dt_t, _ = res[precision]
    e64, g64, dt64, roof64 = res['fp64']
    out.update({'e_shard': e_t, 'e_shard_fp64': e64, 'abs_dE_shard': abs(e_t - e64),
                'rel_dE_shard': abs(e_t - e64) / abs(e64),
                'sec_shard': dt_t, 'sec_shard_fp64': dt64, 'speedup_shard': dt64 / dt_t,
                'sec_fp64_whole_step_extrapolated': dt64 * n_trip / n_shard,
                'grad_rel_err_shard': grad_errors(g_t, g64), 'roofline_fp64': roof64})
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1)
    ap.add_argument('--warmup', type=int, default=0)
    ap.add_argument('--budget', type=float, default=480.0,
                    help='wall-clock seconds from the start of the process; no timed step or '
                         'extra leg starts unless it is expected to finish within it')
    ap.add_argument('--precision', default='bf16x3', choices=['fp64', 'bf16x3', 'fp32x2'])
    ap.add_argument('--compare', default='shard', choices=['shard', 'full', 'none'],
                    help='plain FP64 against the timed form: on one shard of the triples '
                         '(default), on a whole step, or not at all')
    ap.add_argument('--compare-shards', type=int, default=64)
    ap.add_argument('--nocc', type=int, default=40)
    ap.add_argument('--nvir', type=int, default=400)
    ap.add_argument('--naux', type=int, default=None)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--profile-every', type=int, default=64)
    ap.add_argument('--progress-every', type=int, default=512,
                    help='stderr progress line every N occupied triples (0: off)')
    ap.add_argument('--cpu-sample', default='12,120', help='nocc,nvir of the CPU baseline sample')
    ap.add_argument('--cpu-threads', type=int, default=16)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-scaling', default=None, metavar='O,V;O,V;...',
                    help='only time the reference C E(T)+VJP at these sizes on --cpu-threads '
                         'threads and print their GFLOP/s (no GPU work)')
    ap.add_argument('--no-df', action='store_true',
                    help='skip the timing of the DF integral producer and its adjoint')
    ap.add_argument('--energy-only', action='store_true', help='time E(T) without the VJP')
    ap.add_argument('--walk-shards', type=int, default=1,
                    help='time shard 0 of this many round-robin shards of the occupied-triple '
                         'walk instead of the whole step (short runs under hardware counters; '
                         'the line is then not the benchmark metric)')
    ap.add_argument('--out', default=None, help='also write the JSON line to this file')
    ap.add_argument('--calib-copy-mib', type=int, default=0,
                    help='run one plain device copy of this many MiB before the timed region '
                         '(reference dispatch for HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this benchmark runs on one GPU only; the (T) path would shard the '
              'occupied-triple walk round-robin across GPUs (pamd_ccsdt_energy_shard and '
              'pamd_ccsdt_energy_vjp_shard compute one GPU\'s share), which is not run here. '
              f'--gpus {args.gpus} is not supported.', file=sys.stderr)
        sys.exit(2)
    deadline = _T_START + args.budget
    if args.cpu_scaling:
        rows = []
        for item in args.cpu_scaling.split(';'):
            so, sv = (int(x) for x in item.split(','))
            r = cpu_baseline(so, sv, (so, sv), args.cpu_threads)
            if r is None:
                print('bench.py: the reference library (oracle/_ref) is not built', file=sys.stderr)
                sys.exit(1)
            rows.append({'nocc': so, 'nvir': sv, 'naux': 4 * sv, 'sec': r['sample_sec'],
                         'gflops': r['gflops']})
            _log(f'reference C E(T)+VJP o{so}/v{sv}: {r["sample_sec"]:.2f} s, '
                 f'{r["gflops"]:.1f} GFLOP/s')
        model, ncpu = cpu_model()
        line = {'cpu_scaling': rows, 'threads': args.cpu_threads, 'host_cpu': model,
                'host_logical_cpus': ncpu, 'flops': 'E+VJP algorithmic 6 o^3 v^3 (o + v)'}
        print(json.dumps(line), flush=True)
        if args.out:
            with open(args.out, 'w') as f:
                f.write(json.dumps(line) + '\n')
        return

    import torch
    from pyscfad_amd import _lib
    from pyscfad_amd.cc import _ops
    from pyscfad_amd.gen import make_problem

    if not torch.cuda.is_available():
        print('bench.py: no GPU visible', file=sys.stderr)
        sys.exit(1)
    dev = torch.device('cuda', 0)
    o, v = args.nocc, args.nvir
    naux = args.naux or 4 * v
    _log(f'torch and the library loaded; problem o{o}/v{v}/naux{naux}, {args.precision}')

    # untimed warmup: library load + a small problem through every kernel
    tiny = make_problem(3, 40, 160, seed=1)
    tin = _ops.Inputs(tiny.mo_energy, tiny.t1, tiny.t2, tiny.ovvv_packed(), tiny.ovoo,
                      tiny.ovov, None, precision=args.precision, device=dev)
    for _ in range(max(1, args.warmup)):
        if args.energy_only:
            _ops.energy(tin)
        else:
            _ops.energy_vjp(tin)
    torch.cuda.synchronize()
    del tin
    _log(f'warmup: {max(1, args.warmup)} pass(es) on o3/v40')

    t_gen = time.perf_counter()
    prob = make_problem(o, v, naux, seed=args.seed, device=dev)
    idx, idy = torch.tril_indices(v, v, device=dev)
    ovvv_packed = prob.ovvv[:, :, idx, idy].contiguous()
    del prob.ovvv, idx, idy
    torch.cuda.empty_cache()
    inp = _ops.Inputs(prob.mo_energy, prob.t1, prob.t2, ovvv_packed, prob.ovoo, prob.ovov,
                      None, precision=args.precision, device=dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    _log(f'inputs generated on the GPU in {t_gen:.1f} s')
    ws = inp.workspace(not args.energy_only)

    def step():
        if args.walk_shards > 1:
            if args.energy_only:
                return _ops.energy_shard(inp, 0, args.walk_shards), None
            return _ops.energy_vjp_shard(inp, 0, args.walk_shards, et_bar=1.0, workspace=ws)
        if args.energy_only:
            return _ops.energy(inp), None
        return _ops.energy_vjp(inp, et_bar=1.0, workspace=ws)

    if args.calib_copy_mib:
        src = torch.empty(args.calib_copy_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    _lib.profile_reset()
    _lib.profile_enable(args.profile_every)
    # a progress line every ~15 s of work keeps long single-call steps visibly alive
    _lib.profile_progress(args.progress_every)
    step_times = []
    e = grads = None
    # reserve for what follows the timed steps: comparison shard, CPU sample, DF leg
    reserve = 45.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while len(step_times) < max(1, args.steps):
        if step_times and time.time() + max(step_times) + reserve > deadline:
            _log(f'budget: stopping after {len(step_times)} timed step(s) of {args.steps} '
                 f'requested (next would end past {args.budget:.0f} s)')
            break
        ts = time.perf_counter()
        del e, grads
        e, grads = step()
        torch.cuda.synchronize()
        step_times.append(time.perf_counter() - ts)
        _log(f'timed step {len(step_times)}: {step_times[-1]:.2f} s')
    t1 = time.perf_counter()
    _lib.profile_enable(0)
    _lib.profile_progress(0)
    prof = _lib.profile_read()
    nsteps = len(step_times)
    sec = (t1 - t0) / nsteps
    e_val = float(e.item())
    roofline = engine_roofline(prof, args.precision)
    del ws
    torch.cuda.empty_cache()

    F = alg_flops(o, v) if not args.energy_only else alg_flops(o, v) / 3.0
    compare = None
    if args.compare != 'none' and args.precision != 'fp64' and not args.energy_only:
        ref_inp = _ops.Inputs(inp.mo_energy, inp.t1, inp.t2, inp.ovvv, inp.ovoo, inp.ovov, None,
                              precision='fp64', device=dev)
        if args.compare == 'full':
            if time.time() + 2.6 * sec > deadline:
                _log('budget: no time left for a whole FP64 step; comparing on a shard')
                args.compare = 'shard'
        if args.compare == 'shard':
            est = 4.0 * sec / args.compare_shards
            if time.time() + est + 30.0 <= deadline:
                compare = compare_on_shard(inp, ref_inp, args.compare_shards,
                                           max(1, args.profile_every // args.compare_shards),
                                           args.precision)
            else:
                _log('budget: FP64 comparison skipped')
        else:
            _log('FP64 comparison: one whole step on the same inputs')
            _lib.profile_reset()
            _lib.profile_enable(args.profile_every)
            _lib.profile_progress(args.progress_every)
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            e64, g64 = _ops.energy_vjp(ref_inp, et_bar=1.0)
            torch.cuda.synchronize()
            sec64 = time.perf_counter() - c0
            _lib.profile_enable(0)
            _lib.profile_progress(0)
            compare = {'mode': 'full', 'sec_fp64': sec64, 'e_t_fp64': float(e64.item()),
                       'abs_dE': abs(float(e64.item()) - e_val), 'speedup': sec64 / sec,
                       'tflops_alg_fp64': F / sec64 / 1e12,
                       'roofline_fp64': engine_roofline(_lib.profile_read(), 'fp64'),
                       'grad_rel_err': grad_errors(grads, g64)}
            del g64
        del ref_inp
        torch.cuda.empty_cache()
    df = None
    if not args.no_df and not args.energy_only:
        if time.time() + 10.0 < deadline:
            df = df_measure(o, v, naux, args.seed, dev, inp, grads)
            _log('DF producer leg done')
    cpu = None
    if not args.no_cpu_baseline:
        so, sv = (int(x) for x in args.cpu_sample.split(','))
        est = 1.2 * alg_flops(so, sv) / 200e9
        if time.time() + est + 5.0 < deadline:
            try:
                cpu = cpu_baseline(o, v, (so, sv), args.cpu_threads)
            except Exception as exc:  # the compiled reference library may be absent on a host
                cpu = {'value': None, 'unit': 's', 'cores': args.cpu_threads, 'kind': 'reference',
                       'sample': f'not run: {type(exc).__name__}: {exc}'}
            if cpu is not None and args.energy_only and cpu['value']:
                cpu['value'] /= 3.0
            _log('CPU baseline sample done')
        else:
            cpu = {'value': None, 'unit': 's', 'cores': args.cpu_threads, 'kind': 'reference',
                   'sample': 'not run: no time left in the budget'}
    line = {
        'metric': 'E(T)+VJP sec @o40/v400: FP64 MFMA vs split-BF16x3, |dE|<1e-7 Eh',
        'value': sec,
        'unit': 's',
        'n_gpus': 1,
        'steps': nsteps,
        'warmup': args.warmup,
        'steps_requested': args.steps,
        'step_sec': step_times,
        'ms_per_step': sec * 1e3,
        'higher_is_better': False,
        'scaling': 'strong',
        'vs_baseline': None,
        'dtype': 'f64' if args.precision == 'fp64' else args.precision,
        'data': 'synthetic (seeded DF-like generator; see pyscfad_amd/gen.py)',
        'config': {'workload': f'RHF (T) E+VJP nocc={o} nvir={v} naux={naux} seed={args.seed}'
                   + (' (energy only)' if args.energy_only else '')
                   + (f' (walk shard 0 of {args.walk_shards}: not the metric)'
                      if args.walk_shards > 1 else ''),
                   'precision': args.precision, 'ovvv': 'packed',
                   'warmup_problem': 'o3/v40/naux160 (library load, every kernel once)'},
        'e_t': e_val,
        'tflops_alg': F / sec / 1e12,
        'frac_fp64_peak_alg': F / sec / 1e12 / PEAKS['fp64'],
        'input_gen_s': t_gen,
        'roofline': roofline,
        'compare_fp64': compare,
        'df_eris': df,
        'cpu_baseline': cpu,
        'kernels': {k: dict(x, avg_ms=(x['ms'] / x['launches'] if x['launches'] else None))
                    for k, x in prof.items()},
        'wall_s': time.time() - _T_START,
    }
    s = json.dumps(line)
    print(s, flush=True)
    if args.out:
        with open(args.out, 'w') as f:
            f.write(s + '\n')


if __name__ == '__main__':
    main()
