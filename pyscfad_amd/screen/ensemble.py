"""4-member BF16 MLP ensemble predicting pair-resolved (T), with an exact fallback.

``Screen.predict`` returns, per instance, the ensemble-mean E(T) (sum of the
predicted pair contributions) and the spread (standard deviation over the
members of their E(T)).  ``Screen.energy`` accepts the mean when the spread is
at most ``tol`` (1e-5 Eh by default) and otherwise calls the exact kernel.

Inputs and outputs are standardised per column on the training set; widths are
padded to multiples of 8 (the BF16 GEMM's K granularity) with zero columns.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .mlp import MLP, to_bf16

SPREAD_TOL = 1e-5  # Eh


def _pad8(n):
    return (n + 7) // 8 * 8


def _pad_cols(x, n):
    if x.shape[1] == n:
        return x
    return torch.cat([x, x.new_zeros((x.shape[0], n - x.shape[1]))], dim=1)


class Screen:
    def __init__(self, n_in, n_out, device, members=4, hidden=(256, 256), seed=0):
        self.dev = torch.device(device)
        self.n_in, self.n_out = n_in, n_out
        self.p_in, self.p_out = _pad8(n_in), _pad8(n_out)
        self.members = [MLP((self.p_in,) + tuple(hidden) + (self.p_out,), device, seed=seed + 97 * m)
                        for m in range(members)]
        self.x_mean = self.x_std = self.y_mean = self.y_std = None

    def _xin(self, X):
        Xn = ((X - self.x_mean) / self.x_std).float()
        return to_bf16(_pad_cols(Xn, self.p_in).contiguous(), with_t=True)

    def fit(self, X, Y, epochs=300, batch=128, lr=2e-3, wd=1e-5, seed=0, bootstrap=True,
            log=None):
        """X [n, n_in], Y [n, n_out] float64 device tensors (features, pair labels).

        Each member sees its own shuffles (seeded by ``seed`` and its index)
        and, with ``bootstrap``, its own resample (with replacement) of the
        training set, so that the members' disagreement reflects what the data
        leave undetermined; the learning rate follows a cosine schedule.
        Returns the final training loss of each member (standardised units).
        """
        self.x_mean, self.x_std = X.mean(0), X.std(0).clamp_min(1e-30)
        self.y_mean, self.y_std = Y.mean(0), Y.std(0).clamp_min(1e-30)
        Yn = _pad_cols(((Y - self.y_mean) / self.y_std).float(), self.p_out)
        n = X.shape[0]
        batch = min(batch, n // 8 * 8)
        x16_all, _ = self._xin(X)
        hist = []
        for m, net in enumerate(self.members):
            g = torch.Generator(device='cpu').manual_seed(seed * 1000 + m)
            rows = (torch.randint(0, n, (n,), generator=g) if bootstrap
                    else torch.arange(n)).to(self.dev)
            loss = None
            for ep in range(epochs):
                lr_e = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))
                perm = rows[torch.randperm(n, generator=g).to(self.dev)]
                for s0 in range(0, n - batch + 1, batch):
                    idx = perm[s0:s0 + batch]
                    xb = x16_all[idx].contiguous()
                    loss = net.train_step(xb, xb.T.contiguous(), Yn[idx].contiguous(), lr_e, wd)
                if log and (ep % 100 == 0 or ep == epochs - 1):
                    log(f'member {m} epoch {ep} loss {float(loss):.3e}')
            hist.append(float(loss))
        return hist

    def predict_pairs(self, X):
        x16, _ = self._xin(X)
        outs = []
        for net in self.members:
            z = net.forward(x16)[:, :self.n_out].double()
            outs.append(z * self.y_std + self.y_mean)
        return torch.stack(outs)          # [members, n, n_out]

    def predict(self, X):
        e = self.predict_pairs(X).sum(-1)  # [members, n]
        return e.mean(0), e.std(0)

    def energy(self, X, exact_fn, tol=SPREAD_TOL, spread_override=None):
        """E(T) per instance: the ensemble mean where the spread <= tol, else exact_fn(i).

        ``spread_override`` (tensor of spreads) replaces the computed spreads,
        which lets a caller force either branch.  Returns (energies, spreads,
        fallback indices, seconds spent in exact_fn).
        """
        mean, spread = self.predict(X)
        if spread_override is not None:
            spread = torch.as_tensor(spread_override, dtype=mean.dtype, device=mean.device)
        out = mean.clone()
        fallback = (spread > tol).nonzero().flatten().tolist()
        t0 = time.perf_counter()
        for i in fallback:
            out[i] = exact_fn(i)
        return out, spread, fallback, time.perf_counter() - t0
