"""The closed-shell (T) oracle against an independent spin-orbital statement
(oracle/gccsd_t_oracle.py, after pyscfad/cc/gccsd_t.py): the total E(T), and
the occupied-orbital partition -- a projector on one occupied orbital in the
spin-orbital form equals the row sum of the closed-shell pair energies, i.e.
the reference's occupied-weighted (T) with mat = e_i e_i^T."""
import numpy as np
import pytest

from oracle import ccsd_t_oracle as R
from oracle import gccsd_t_oracle as G


def _problem(o, v, naux, seed, fvo_scale):
    n = o + v
    rng = np.random.default_rng(seed)
    g = rng.normal(0, 0.3 / np.sqrt(naux), (naux, n, n))
    L = np.triu(g) + np.swapaxes(np.triu(g, 1), 1, 2)
    mo = np.concatenate([np.sort(rng.uniform(-1.5, -0.3, o)), np.sort(rng.uniform(0.1, 3.0, v))])
    fock = np.diag(mo)
    fvo = rng.normal(0, fvo_scale, (v, o)) if fvo_scale else np.zeros((v, o))
    fock[o:, :o] = fvo
    fock[:o, o:] = fvo.T
    Lov = L[:, :o, o:].reshape(naux, -1)
    ovvv = (Lov.T @ L[:, o:, o:].reshape(naux, -1)).reshape(o, v, v, v)
    ovoo = (Lov.T @ L[:, :o, :o].reshape(naux, -1)).reshape(o, v, o, o)
    ovov = (Lov.T @ Lov).reshape(o, v, o, v)
    eia = mo[:o, None] - mo[None, o:]
    t2 = ovov.transpose(0, 2, 1, 3) / (eia[:, None, :, None] + eia[None, :, None, :])
    t1 = rng.normal(0, 1e-2, (o, v))
    return L, mo, fock, fvo, t1, t2, ovvv, ovoo, ovov


@pytest.mark.parametrize('o,v,naux,seed,fvo_scale', [(2, 4, 10, 0, 0.0), (3, 5, 12, 1, 1e-2),
                                                     (4, 6, 14, 2, 0.0)])
def test_spin_orbital_and_closed_shell_agree(o, v, naux, seed, fvo_scale):
    L, mo, fock, fvo, t1, t2, ovvv, ovoo, ovov = _problem(o, v, naux, seed, fvo_scale)
    e_rhf = R.energy(mo, t1, t2, ovvv, ovoo, ovov, fvo)
    e_so = G.energy(L, mo, fock, t1, t2, o)
    assert e_rhf != 0.0
    assert abs(e_rhf - e_so) <= 1e-13 * abs(e_rhf)
    pairs = R.pair_energy(mo, t1, t2, ovvv, ovoo, ovov, fvo)
    for i in range(o):
        e_i = G.energy(L, mo, fock, t1, t2, o, prj=np.eye(o)[i:i + 1])
        assert abs(e_i - pairs[i].sum()) <= 1e-12 * abs(e_rhf)


def test_spin_orbital_form_sees_a_planted_error():
    """Dropping the exchange-signed mixed-spin blocks of t2 changes E(T)."""
    L, mo, fock, fvo, t1, t2, ovvv, ovoo, ovov = _problem(2, 4, 10, 0, 0.0)
    e_rhf = R.energy(mo, t1, t2, ovvv, ovoo, ovov, fvo)
    good = G.spin_t2
    try:
        def bad(t):
            s = good(t)
            s[0::2, 1::2, 1::2, 0::2] = 0.0
            return s
        G.spin_t2 = bad
        e_bad = G.energy(L, mo, fock, t1, t2, 2)
    finally:
        G.spin_t2 = good
    assert abs(e_bad - e_rhf) > 1e-6 * abs(e_rhf)
