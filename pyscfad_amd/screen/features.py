"""MP2-level pair features and pair-resolved (T) labels.

Features (128 for nocc = 16): the 120 unordered occupied-pair MP2 energies
eps_ij = e_ij + e_ji (i < j), with
e_ij = sum_ab (2 t2[i,j,a,b] - t2[j,i,a,b]) (ia|jb) and the MP2 amplitudes
t2[i,j,a,b] = (ia|jb) / (e_i + e_j - e_a - e_b) -- the pair resolution of the
MP2 energy of pyscfad/cc/ccsd.py:376-381, computed by the HIP kernel behind
``pamd_mp2_init_amps`` -- followed by 8 orbital statistics: HOMO-LUMO gap,
mean/std/min of the occupied energies, mean/std/max of the virtual energies,
and the total MP2 energy.

Labels (136 for nocc = 16): the pair-resolved (T) energy of the exact kernel
(``pamd_ccsdt_pair_energy``), folded over unordered pairs i <= j:
tau_ij = e(T)_ij + e(T)_ji (i < j), tau_ii = e(T)_ii; their sum is E(T).
"""
from __future__ import annotations

import torch

N_GAP = 8


def n_features(nocc):
    return nocc * (nocc - 1) // 2 + N_GAP


def n_labels(nocc):
    return nocc * (nocc + 1) // 2


def features_from_pairs(mo_energy, e, nocc):
    """Feature vector from the (o, o) MP2 pair energies ``e`` (host logic)."""
    iu, ju = torch.triu_indices(nocc, nocc, offset=1, device=e.device)
    pair = (e + e.T)[iu, ju]
    eo, ev = mo_energy[:nocc], mo_energy[nocc:]
    stats = torch.stack([ev.min() - eo.max(), eo.mean(), eo.std(), eo.min(),
                         ev.mean(), ev.std(), ev.max(), e.sum()])
    return torch.cat([pair, stats])


def features(mo_energy, ovov, nocc):
    """Feature vector of one instance (MP2 pair energies on the GPU)."""
    from ..cc.mp2 import pair_energies
    return features_from_pairs(mo_energy, pair_energies(mo_energy, ovov, nocc), nocc)


def fold_pairs(e_pair):
    """(o, o) pair-resolved energies -> i <= j vector (sum preserved)."""
    o = e_pair.shape[0]
    iu, ju = torch.triu_indices(o, o, offset=0, device=e_pair.device)
    # a where() rather than a masked assignment: masks need a host-side count
    return torch.where(iu == ju, e_pair[iu, iu], (e_pair + e_pair.T)[iu, ju])
