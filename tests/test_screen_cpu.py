"""Learned-screen host logic on the CPU: features, label folding, data split,
instance generation (no kernel calls)."""
import numpy as np
import torch

from pyscfad_amd.gen import make_problem
from pyscfad_amd.screen import data as D
from pyscfad_amd.screen.features import fold_pairs, n_features, n_labels


def test_sizes_of_the_screen_workload():
    assert n_features(16) == 128 and n_labels(16) == 136
    tr, te = D.split()
    assert len(tr) == 1800 and len(te) == 248
    assert len(set(tr) | set(te)) == 2048


def test_fold_pairs_preserves_the_sum():
    e = torch.randn(6, 6, generator=torch.Generator().manual_seed(0), dtype=torch.float64)
    y = fold_pairs(e)
    assert y.shape == (21,)
    assert torch.isclose(y.sum(), e.sum(), rtol=1e-14)
    assert y[0] == e[0, 0] and y[1] == e[0, 1] + e[1, 0]


def test_instances_are_seeded_and_differ():
    iset = D.InstanceSet('cpu', 3, 10, 30)
    a, b, a2 = iset.instance(1), iset.instance(2), iset.instance(1)
    assert torch.equal(a.ovvv, a2.ovvv) and torch.equal(a.mo_energy, a2.mo_energy)
    assert not torch.equal(a.ovvv, b.ovvv)
    base = make_problem(3, 10, 30, seed=0)
    rel = float(np.linalg.norm(a.ovov.numpy() - base.ovov) / np.linalg.norm(base.ovov))
    assert 0 < rel < 0.1
    assert torch.all(a.mo_energy[:3] < 0) and torch.all(torch.diff(a.mo_energy[3:]) >= 0)
