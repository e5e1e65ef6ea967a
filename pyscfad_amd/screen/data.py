"""Perturbed (T) instances for the learned screen (nocc=16, nvir=160, naux=640).

Instance s (s = 1..2048) is the seed-0 base problem of :mod:`pyscfad_amd.gen` with
* the upper triangle of L displaced by 0.02 * xi_s * sigma_L * z, where
  xi_s ~ U[-1, 1] is drawn from ``numpy.random.default_rng(10000 + s)``,
  sigma_L = 0.1/sqrt(naux) is the scale of L itself (so the displacement is a
  relative one, like a scan step), and z ~ N(0, 1) comes from torch's device
  generator seeded with 20000 + s (drawn on the GPU: an instance costs no
  host-side random numbers beyond a few hundred);
* every orbital energy shifted by N(0, 0.05) drawn from the same numpy
  generator after xi_s (occupied first, then virtual; each block re-sorted);
* t1 and the base draws unchanged.
Labels: the exact FP64 kernel's pair-resolved (T), folded to i <= j.
Features: MP2 pair energies and orbital statistics (:mod:`.features`).
The train/test split is the first 1800 / last 248 of a permutation of the
instance list drawn from ``numpy.random.default_rng(7)``.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import gen
from ..cc import _ops
from .features import features, fold_pairs

NOCC, NVIR, NAUX = 16, 160, 640
N_INSTANCES, N_TRAIN = 2048, 1800
L_REL, E_SIGMA = 0.02, 0.05


def split(n=N_INSTANCES, n_train=N_TRAIN, seed=7):
    perm = np.random.default_rng(seed).permutation(n)
    return perm[:n_train], perm[n_train:]


class InstanceSet:
    """Base draws cached on the device; instances formed on demand."""

    def __init__(self, device, nocc=NOCC, nvir=NVIR, naux=NAUX):
        self.dev = torch.device(device)
        self.nocc, self.nvir, self.naux = nocc, nvir, naux
        (iu, ju, tri), self.e_occ, self.e_vir, t1, _ = gen._draw(nocc, nvir, naux, 0)
        self.iu = torch.from_numpy(iu).to(self.dev)
        self.ju = torch.from_numpy(ju).to(self.dev)
        self.tri = torch.from_numpy(tri).to(self.dev)
        self.t1 = torch.from_numpy(t1).to(self.dev)
        self.sigma = 0.1 / np.sqrt(naux)

    def instance(self, s):
        o, v = self.nocc, self.nvir
        rng = np.random.default_rng(10_000 + s)
        xi = rng.uniform(-1.0, 1.0)
        e_occ = np.sort(self.e_occ + rng.normal(0.0, E_SIGMA, size=o))
        e_vir = np.sort(self.e_vir + rng.normal(0.0, E_SIGMA, size=v))
        g = torch.Generator(device=self.dev).manual_seed(20_000 + s)
        z = torch.randn(self.tri.shape, generator=g, dtype=torch.float64, device=self.dev)
        tri = self.tri + (L_REL * xi * self.sigma) * z
        del z
        L = gen.unpack_L_torch(self.iu, self.ju, tri, self.naux, o + v)
        # the orbital energies reach the device through pinned memory without
        # blocking the host (a pageable copy would wait for the stream)
        mo = torch.from_numpy(np.concatenate([e_occ, e_vir]))
        if self.dev.type == 'cuda':
            mo = mo.pin_memory().to(self.dev, non_blocking=True)
        return gen.assemble_torch(L, mo, torch.diag(mo), self.t1, o, v)


def exact_pairs(p, precision='fp64'):
    inp = _ops.Inputs(p.mo_energy, p.t1, p.t2, p.ovvv, p.ovoo, p.ovov, None,
                      precision=precision, device=p.t2.device)
    return _ops.pair_energy(inp)


def build(seeds, device, log=None, every=128, streams=4):
    """(X [n, 128], Y [n, 136], E [n]) float64 device tensors for instance ``seeds``.

    Instances are formed and labelled round-robin on ``streams`` HIP streams (at
    the default priority): one instance's walk is a chain of small launches
    (about 600 workgroups per W-build GEMM at o16/v160, under two rounds of the
    GPU's workgroup slots), so several instances in flight fill the GPU.  Every
    tensor of an instance is allocated on its own stream, so the caching
    allocator never hands its memory to another stream while it is in use."""
    iset = InstanceSet(device)
    dev = iset.dev
    sts = [torch.cuda.Stream(dev) for _ in range(max(1, streams))]
    cur = torch.cuda.current_stream(dev)
    for st in sts:
        st.wait_stream(cur)   # the cached base draws are ready
    n = len(seeds)
    X, Y, E = [None] * n, [None] * n, [None] * n
    for c, s in enumerate(seeds):
        with torch.cuda.stream(sts[c % len(sts)]):
            p = iset.instance(int(s))
            ep = exact_pairs(p)
            X[c] = features(p.mo_energy, p.ovov, p.nocc)
            Y[c] = fold_pairs(ep)
            E[c] = ep.sum()
            del p, ep
        if log and (c % every == 0 or c == n - 1):
            log(f'instance {c + 1}/{n} seed {int(s)} queued')
    for st in sts:
        cur.wait_stream(st)
    # the per-instance results are read on the caller's stream: keep their
    # memory from being reused on their own streams before that read is done
    for t in X + Y + E:
        t.record_stream(cur)
    return torch.stack(X), torch.stack(Y), torch.stack(E)
