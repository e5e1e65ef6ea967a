"""The learned-screen workload: 2048 perturbed o16/v160/naux640 instances.

    python -m pyscfad_amd.screen.experiment [--n 2048] [--streams 4] [--data PATH]
                                            [--save-data PATH] [--out report.json]

Labels come from the exact FP64 pair-resolved (T) kernel; the 4-member BF16
ensemble (128 -> 256 -> 256 -> 136) is trained by the library's kernels on the
first 1800 instances of the seeded split and evaluated on the other 248.  The
report gives the test errors, the fraction the 1e-5 Eh spread gate accepts,
the error of the accepted answers, the same figures for the estimate available
without the network (E(T) = c E(MP2), c fitted on the training set), and the
time of the screened pass (inference + exact fallbacks) against the exact pass.
The labels run on several HIP streams at once (``--streams``); the exact cost
per instance is the label throughput's reciprocal.  A progress line is
printed every 256 queued instances and every 100 epochs.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import data as D
from .ensemble import SPREAD_TOL, Screen
from .features import features, fold_pairs


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--n', type=int, default=D.N_INSTANCES)
    ap.add_argument('--n-train', type=int, default=D.N_TRAIN)
    ap.add_argument('--data', default=None, help='load (X, Y, E, t_exact) from this .npz')
    ap.add_argument('--save-data', default=None)
    ap.add_argument('--epochs', type=int, default=600)
    ap.add_argument('--batch', type=int, default=128)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--no-bootstrap', action='store_true')
    ap.add_argument('--streams', type=int, default=4,
                    help='HIP streams the labels run on concurrently')
    ap.add_argument('--labels-only', action='store_true',
                    help='stop after the labels (throughput measurement)')
    ap.add_argument('--out', default=None)
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        log('experiment: no GPU visible')
        return 1
    dev = torch.device('cuda', 0)

    if args.data:
        z = np.load(args.data)
        X, Y, E = (torch.from_numpy(z[k]).to(dev) for k in ('X', 'Y', 'E'))
        t_exact = z['t_exact']
    else:
        torch.cuda.synchronize()
        t_all = time.perf_counter()
        X, Y, E = D.build(range(1, args.n + 1), dev, log=log, every=256, streams=args.streams)
        torch.cuda.synchronize()
        t_all = time.perf_counter() - t_all
        # labels run on several streams at once: the cost of one exact (T) is the
        # label throughput's reciprocal
        t_exact = np.full(args.n, t_all / args.n)
        log(f'labels: {args.n} instances in {t_all:.1f} s on {args.streams} streams, '
            f'{t_all / args.n * 1e3:.1f} ms/instance')
        if args.save_data:
            np.savez(args.save_data, X=X.cpu().numpy(), Y=Y.cpu().numpy(), E=E.cpu().numpy(),
                     t_exact=t_exact)

    if args.labels_only:
        return 0
    tr, te = D.split(X.shape[0], args.n_train)
    tr_d, te_d = torch.as_tensor(tr, device=dev), torch.as_tensor(te, device=dev)
    sc = Screen(X.shape[1], Y.shape[1], dev, members=4)
    t0 = time.perf_counter()
    losses = sc.fit(X[tr_d], Y[tr_d], epochs=args.epochs, batch=args.batch, lr=args.lr,
                    bootstrap=not args.no_bootstrap, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mean, spread = sc.predict(X[te_d])
    torch.cuda.synchronize()
    t_infer = time.perf_counter() - t0
    err = (mean - E[te_d]).abs()
    emp2 = X[:, -1]
    c = float((emp2[tr_d] * E[tr_d]).sum() / (emp2[tr_d] ** 2).sum())
    err_mp2 = (c * emp2[te_d] - E[te_d]).abs()
    acc = spread <= SPREAD_TOL
    n_te = len(te)
    exact_pass = float(t_exact[te].sum())
    screened_pass = t_infer + float(t_exact[te][~acc.cpu().numpy()].sum())
    tr_mean, tr_spread = sc.predict(X[tr_d])
    rep = {
        'workload': 'learned screen: o16/v160/naux640, 2048 perturbed instances, 1800/248 split',
        'n_instances': int(X.shape[0]), 'n_train': len(tr), 'n_test': n_te,
        'features': int(X.shape[1]), 'labels': int(Y.shape[1]),
        'ensemble': '4 x MLP 128-256-256-136, SiLU, bf16 weights/activations, fp32 accumulation',
        'e_t_mean': float(E.mean()), 'e_t_std': float(E.std()),
        'train_loss_std_units': losses, 'train_s': t_train,
        'train_mae': float((tr_mean - E[tr_d]).abs().mean()),
        'test_mae': float(err.mean()), 'test_max_err': float(err.max()),
        'test_rmse': float(err.pow(2).mean().sqrt()),
        'mp2_scaling_c': c, 'mp2_scaling_test_mae': float(err_mp2.mean()),
        'mp2_scaling_test_max_err': float(err_mp2.max()),
        'spread_tol': SPREAD_TOL, 'accepted_fraction': float(acc.double().mean()),
        'accepted_max_err': float(err[acc].max()) if bool(acc.any()) else None,
        'accepted_mae': float(err[acc].mean()) if bool(acc.any()) else None,
        'rejected_max_err': float(err[~acc].max()) if bool((~acc).any()) else None,
        'spread_median': float(spread.median()),
        'bootstrap': not args.no_bootstrap, 'epochs': args.epochs,
        'err_over_spread_median': float((err / spread.clamp_min(1e-30)).median()),
        'accepted_err_gt_tol': int(((err > SPREAD_TOL) & acc).sum()),
        'exact_ms_per_instance': float(t_exact.mean() * 1e3),
        'infer_ms_test_set': t_infer * 1e3,
        'exact_pass_s': exact_pass, 'screened_pass_s': screened_pass,
        'speedup': exact_pass / screened_pass if screened_pass > 0 else None,
    }
    log(json.dumps(rep))
    if args.out:
        with open(args.out, 'w') as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == '__main__':
    sys.exit(main())
