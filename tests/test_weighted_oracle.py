"""The occupied-weighted (T) of a diagonal projector in the oracle: energy
against the pair energies and the spin-orbital projector form, adjoint
against central differences, and m = 1 against the plain adjoint."""
import numpy as np

from oracle import ccsd_t_oracle as O
from oracle import gccsd_t_oracle as G
from pyscfad_amd.gen import make_L, make_problem

BLOCKS = ('t1', 't2', 'ovvv', 'ovoo', 'ovov', 'fvo')


def _args(p):
    return (p.mo_energy, p.t1, p.t2, p.ovvv, p.ovoo, p.ovov, p.fvo)


def test_weighted_energy_is_the_weighted_pair_row_sum():
    p = make_problem(4, 9, 30, seed=5, fvo_scale=1e-3)
    m = np.array([0.3, 1.0, 0.0, 0.7])
    e_m, _ = O.energy_vjp(*_args(p), occ_weight=m)
    pairs = O.pair_energy(*_args(p))
    assert np.isclose(e_m, (m * pairs.sum(1)).sum(), rtol=1e-13)
    # the spin-orbital form with the projector rows sqrt(m_i) e_i
    L = make_L(4, 9, 30, seed=5)
    e_so = G.energy(L, p.mo_energy, p.fock, p.t1, p.t2, 4, prj=np.diag(np.sqrt(m)))
    assert abs(e_so - e_m) <= 1e-12 * abs(e_m)


def test_unit_weights_give_the_plain_adjoint():
    p = make_problem(3, 8, 24, seed=6)
    e1, g1 = O.energy_vjp(*_args(p))
    e2, g2 = O.energy_vjp(*_args(p), occ_weight=np.ones(3))
    assert np.isclose(e1, e2, rtol=1e-14)
    for k in g1:
        np.testing.assert_allclose(g2[k], g1[k], rtol=1e-13, atol=1e-18)


def test_weighted_adjoint_central_differences():
    p = make_problem(3, 7, 20, seed=8, fvo_scale=1e-3)
    m = np.array([0.2, 0.9, 0.5])
    _e, g = O.energy_vjp(*_args(p), occ_weight=m)
    grads = dict(t1=g['t1_bar'], t2=g['t2_bar'], ovvv=g['ovvv_bar'], ovoo=g['ovoo_bar'],
                 ovov=g['ovov_bar'], fvo=g['fvo_bar'])
    rng = np.random.default_rng(1)
    x = dict(t1=p.t1, t2=p.t2, ovvv=p.ovvv, ovoo=p.ovoo, ovov=p.ovov, fvo=p.fvo)
    h = 1e-5
    for _ in range(3):
        d = {k: rng.normal(size=x[k].shape) for k in BLOCKS}
        nrm = np.sqrt(sum((t * t).sum() for t in d.values()))
        d = {k: t / nrm for k, t in d.items()}

        def em(s):
            y = {k: x[k] + s * h * d[k] for k in BLOCKS}
            return O.energy_vjp(p.mo_energy, y['t1'], y['t2'], y['ovvv'], y['ovoo'], y['ovov'],
                                y['fvo'], occ_weight=m)[0]
        fd = (em(1) - em(-1)) / (2 * h)
        ana = sum((grads[k] * d[k]).sum() for k in BLOCKS)
        assert abs(fd - ana) <= 1e-6 * abs(fd), (fd, ana)
    # mo_energy: the derivative along each orbital energy
    for q in (0, 4):
        e = p.mo_energy.copy()
        e[q] += h
        ep = O.energy_vjp(e, *_args(p)[1:], occ_weight=m)[0]
        e[q] -= 2 * h
        emn = O.energy_vjp(e, *_args(p)[1:], occ_weight=m)[0]
        assert abs((ep - emn) / (2 * h) - g['mo_energy_bar'][q]) <= 1e-6 * abs(g['mo_energy_bar'][q])


def test_projected_energy_matches_spin_orbital_projector_and_limits():
    p = make_problem(4, 8, 28, seed=9, fvo_scale=1e-3)
    L = make_L(4, 8, 28, seed=9)
    rng = np.random.default_rng(4)
    ulo = rng.normal(size=(2, 4))
    e = O.projected_energy(*_args(p), ulo)
    e_so = G.energy(L, p.mo_energy, p.fock, p.t1, p.t2, 4, prj=ulo)
    assert abs(e - e_so) <= 1e-12 * abs(e_so)
    # identity projector -> E(T); one orbital -> pair-energy row sum
    e_t = O.energy(*_args(p))
    assert np.isclose(O.projected_energy(*_args(p), np.eye(4)), e_t, rtol=1e-13)
    pairs = O.pair_energy(*_args(p))
    assert np.isclose(O.projected_energy(*_args(p), np.eye(4)[1:2]), pairs[1].sum(), rtol=1e-12)
    # a non-diagonal projector is not the diagonal-weighted energy of its diagonal
    m = (ulo.T @ ulo).diagonal()
    e_diag, _ = O.energy_vjp(*_args(p), occ_weight=m)
    assert abs(e - e_diag) > 1e-6 * abs(e)


def test_projected_adjoint_central_differences_and_identity():
    p = make_problem(3, 7, 20, seed=12, fvo_scale=1e-3)
    rng = np.random.default_rng(5)
    ulo = rng.normal(size=(2, 3))
    e, g = O.projected_energy_vjp(*_args(p), ulo, et_bar=1.0)
    assert np.isclose(e, O.projected_energy(*_args(p), ulo), rtol=1e-14)
    x = dict(t1=p.t1, t2=p.t2, ovvv=p.ovvv, ovoo=p.ovoo, ovov=p.ovov, fvo=p.fvo)
    grads = dict(t1=g['t1_bar'], t2=g['t2_bar'], ovvv=g['ovvv_bar'], ovoo=g['ovoo_bar'],
                 ovov=g['ovov_bar'], fvo=g['fvo_bar'])
    h = 1e-5
    for _ in range(3):
        d = {k: rng.normal(size=x[k].shape) for k in BLOCKS}
        nrm = np.sqrt(sum((t * t).sum() for t in d.values()))
        d = {k: t / nrm for k, t in d.items()}

        def ep(s):
            y = {k: x[k] + s * h * d[k] for k in BLOCKS}
            return O.projected_energy(p.mo_energy, y['t1'], y['t2'], y['ovvv'], y['ovoo'],
                                      y['ovov'], y['fvo'], ulo)
        fd = (ep(1) - ep(-1)) / (2 * h)
        ana = sum((grads[k] * d[k]).sum() for k in BLOCKS)
        assert abs(fd - ana) <= 1e-6 * abs(fd), (fd, ana)
    for q in (1, 5):
        em = p.mo_energy.copy()
        em[q] += h
        a = O.projected_energy(em, *_args(p)[1:], ulo)
        em[q] -= 2 * h
        b = O.projected_energy(em, *_args(p)[1:], ulo)
        assert abs((a - b) / (2 * h) - g['mo_energy_bar'][q]) <= 1e-6 * abs(g['mo_energy_bar'][q])
    for f, l in ((0, 0), (1, 2)):
        u = ulo.copy()
        u[f, l] += h
        a = O.projected_energy(*_args(p), u)
        u[f, l] -= 2 * h
        b = O.projected_energy(*_args(p), u)
        assert abs((a - b) / (2 * h) - g['ulo_bar'][f, l]) <= 1e-6 * abs(g['ulo_bar'][f, l])
    # ulo = I: the plain adjoint
    _e0, g0 = O.energy_vjp(*_args(p), et_bar=1.0)
    _e1, g1 = O.projected_energy_vjp(*_args(p), np.eye(3), et_bar=1.0)
    for k in g0:
        np.testing.assert_allclose(g1[k], g0[k], rtol=1e-12, atol=1e-18)


def test_projected_adjoint_matches_reference_lno_vjp_golden():
    """The oracle's projected (T) and its adjoint, including the full mat_bar,
    against the golden vectors of the reference's own lnoccsdt_contract /
    lnoccsdt_energy_vjp (pyscfadlib/.../vjp/cc/lno_ccsd_t_vjp.c:400-566, driven as
    pyscfad/lno/ccsd_t.py:124-237) for a non-diagonal mat = ulo^T ulo."""
    import os
    from .conftest import GOLDEN
    for name in ('o4v12', 'o6v20_fvo'):
        z = np.load(os.path.join(GOLDEN, f'lno_ccsd_t_{name}.npz'))
        o, v = int(z['nocc']), int(z['nvir'])
        p = make_problem(o, v, int(z['naux']), seed=int(z['seed']),
                         fvo_scale=float(z['fvo_scale']))
        ulo = z['ulo']
        assert np.count_nonzero(ulo.T @ ulo - np.diag(np.diag(ulo.T @ ulo))) > 0
        e, g = O.projected_energy_vjp(*_args(p), ulo, et_bar=float(z['et_bar']))
        assert abs(e - float(z['e_t'])) <= 1e-13 * abs(float(z['e_t']))
        for k in ('t1_bar', 't2_bar', 'fvo_bar', 'mo_energy_bar', 'ovoo_bar', 'ovov_bar',
                  'mat_bar', 'ulo_bar'):
            ref = z[k]
            assert np.linalg.norm(g[k] - ref) <= 1e-12 * np.linalg.norm(ref), (name, k)
        packed = O.fold_ovvv_packed(g['ovvv_bar'])
        assert (np.linalg.norm(packed - z['ovvv_bar_packed'])
                <= 1e-12 * np.linalg.norm(z['ovvv_bar_packed'])), name
        # the chain rule of mat = ulo^T ulo, on the reference's own mat_bar
        assert np.allclose(z['ulo_bar'], ulo @ (z['mat_bar'] + z['mat_bar'].T), rtol=1e-14)
