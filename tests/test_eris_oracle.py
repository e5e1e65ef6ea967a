"""Pins of the MP2 and DF-integral oracle (oracle/eris_oracle.py) on the CPU,
and the host logic of the screen features built on the MP2 pair energies."""
import numpy as np
import torch

from oracle import eris_oracle as EO
from pyscfad_amd.gen import make_problem
from pyscfad_amd.screen.features import features_from_pairs, n_features


def _symmetric_L(naux, nmo, seed):
    rng = np.random.default_rng(seed)
    G = rng.normal(0, 0.3, (naux, nmo, nmo))
    return np.triu(G) + np.swapaxes(np.triu(G, 1), 1, 2)


def test_mp2_oracle_matches_loops_and_the_reference_emp2():
    p = make_problem(3, 7, 20, seed=4)
    emp2, t2, e = EO.mp2_init_amps(p.mo_energy, p.ovov, 3)
    t2_l, e_l = EO.mp2_pairs_loops(p.mo_energy, p.ovov, 3)
    assert np.array_equal(t2, t2_l)          # same denominator order -> bitwise
    assert np.allclose(e, e_l, rtol=1e-13, atol=0)
    # the two einsums of pyscfad/cc/ccsd.py:380-381
    ref = 2 * np.einsum('ijab,iajb', t2, p.ovov) - np.einsum('jiab,iajb', t2, p.ovov)
    assert np.isclose(emp2, ref, rtol=1e-14)
    # the generator's amplitudes are the MP2 amplitudes
    assert np.allclose(t2, p.t2, rtol=1e-15, atol=0)


def test_mp2_oracle_closed_form_one_occupied_orbital():
    rng = np.random.default_rng(3)
    v = 5
    ovov = rng.normal(size=(1, v, 1, v))
    mo = np.concatenate([[-0.7], np.sort(rng.uniform(0.2, 1.0, v))])
    _e, _t2, pairs = EO.mp2_init_amps(mo, ovov, 1)
    D = (mo[0] - mo[1:])[:, None] + (mo[0] - mo[1:])[None, :]
    assert np.isclose(pairs[0, 0], (ovov[0, :, 0, :] ** 2 / D).sum(), rtol=1e-14)


def test_mp2_oracle_breaks_on_a_planted_sign_error():
    """A wrong exchange sign changes the pair energies by far more than rounding."""
    p = make_problem(3, 7, 20, seed=4)
    _e, t2, e = EO.mp2_init_amps(p.mo_energy, p.ovov, 3)
    wrong = 2 * np.einsum('ijab,iajb->ij', t2, p.ovov) + np.einsum('jiab,iajb->ij', t2, p.ovov)
    assert np.abs(wrong - e).max() > 1e3 * np.abs(e).max() * 1e-13


def test_df_oracle_matches_loops():
    L = _symmetric_L(4, 6, seed=1)
    got = EO.df_eris(L, 2)
    ref = EO.df_eris_loops(L, 2)
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        assert np.allclose(g, r, rtol=1e-13, atol=1e-15)


def test_df_oracle_adjoint_identity_and_lower_triangle():
    rng = np.random.default_rng(2)
    o, nmo, naux = 3, 8, 5
    L = _symmetric_L(naux, nmo, seed=2)
    v = nmo - o
    bars = [rng.normal(size=s) for s in ((o, v, v * (v + 1) // 2), (o, v, o, o), (o, v, o, v))]
    Lb = EO.df_eris_vjp(L, o, *bars)
    for _ in range(3):
        dL = rng.normal(size=L.shape)
        lhs = (Lb * dL).sum()
        # the blocks are quadratic in L, so the central difference is exact
        # up to rounding
        h = 1e-3
        plus = EO.df_eris(L + h * dL, o)
        minus = EO.df_eris(L - h * dL, o)
        rhs = sum(((p_ - m_) / (2 * h) * b).sum() for p_, m_, b in zip(plus, minus, bars))
        assert np.isclose(lhs, rhs, rtol=1e-10), (lhs, rhs)
    # pack_tril's adjoint leaves the upper virtual triangle and the vo block zero
    iu = np.triu_indices(v, 1)
    assert np.all(Lb[:, o:, o:][:, iu[0], iu[1]] == 0)
    assert np.all(Lb[:, o:, :o] == 0)


def test_df_oracle_adjoint_matches_torch_autograd():
    o, nmo, naux = 2, 7, 4
    L = _symmetric_L(naux, nmo, seed=5)
    v = nmo - o
    rng = np.random.default_rng(6)
    bars = [rng.normal(size=s) for s in ((o, v, v * (v + 1) // 2), (o, v, o, o), (o, v, o, v))]
    Lt = torch.tensor(L, requires_grad=True)
    Loo = Lt[:, :o, :o].reshape(naux, -1)
    Lov = Lt[:, :o, o:].reshape(naux, -1)
    b, c = np.tril_indices(v)
    Lvv = Lt[:, o:, o:][:, b, c]
    outs = [(Lov.T @ Lvv).reshape(o, v, -1), (Lov.T @ Loo).reshape(o, v, o, o),
            (Lov.T @ Lov).reshape(o, v, o, v)]
    s = sum((x * torch.from_numpy(g)).sum() for x, g in zip(outs, bars))
    s.backward()
    assert np.allclose(EO.df_eris_vjp(L, o, *bars), Lt.grad.numpy(), rtol=1e-12, atol=1e-14)
    # a missing cotangent counts as zero
    only = EO.df_eris_vjp(L, o, None, bars[1], None)
    assert np.allclose(only, EO.df_eris_vjp(L, o, np.zeros_like(bars[0]), bars[1],
                                            np.zeros_like(bars[2])))


def test_features_from_pairs_host_logic():
    p = make_problem(4, 12, 40, seed=2)
    _e, _t2, e = EO.mp2_init_amps(p.mo_energy, p.ovov, 4)
    f = features_from_pairs(torch.from_numpy(p.mo_energy), torch.from_numpy(e), 4).numpy()
    assert f.shape == (n_features(4),)
    iu, ju = np.triu_indices(4, 1)
    assert np.allclose(f[:6], (e + e.T)[iu, ju])
    eo, ev = p.mo_energy[:4], p.mo_energy[4:]
    assert np.isclose(f[6], ev.min() - eo.max())
    assert np.isclose(f[-1], e.sum(), rtol=1e-14)
