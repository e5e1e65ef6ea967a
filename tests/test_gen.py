"""The seeded generator: determinism, symmetry and denominator bounds."""
import numpy as np

from pyscfad_amd.gen import make_problem


def test_deterministic_and_symmetric():
    a = make_problem(4, 12, 48, seed=3)
    b = make_problem(4, 12, 48, seed=3)
    assert np.array_equal(a.ovvv, b.ovvv) and np.array_equal(a.t1, b.t1)
    # (ia|bc) = (ia|cb), (ia|jb) = (jb|ia), t2[i,j,a,b] = t2[j,i,b,a]
    assert np.allclose(a.ovvv, a.ovvv.transpose(0, 1, 3, 2), atol=0, rtol=1e-14)
    assert np.allclose(a.ovov, a.ovov.transpose(2, 3, 0, 1), atol=0, rtol=1e-14)
    assert np.allclose(a.t2, a.t2.transpose(1, 0, 3, 2), atol=0, rtol=1e-14)
    assert np.allclose(a.ovoo, a.ovoo.transpose(0, 1, 3, 2), atol=0, rtol=1e-14)


def test_energies_and_denominators():
    p = make_problem(6, 20, 80, seed=0)
    eo, ev = p.mo_energy[:6], p.mo_energy[6:]
    assert np.all(np.diff(eo) >= 0) and np.all(np.diff(ev) >= 0)
    assert eo.min() >= -1.5 and eo.max() <= -0.3 and ev.min() >= 0.1 and ev.max() <= 3.0
    assert np.all(p.fvo == 0)
    # MP2 amplitudes from the ERIs
    d = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    assert np.allclose(p.t2 * d, p.ovov.transpose(0, 2, 1, 3), rtol=1e-13, atol=1e-20)


def test_fvo_and_packing():
    p = make_problem(3, 9, 36, seed=1, fvo_scale=1e-3)
    assert np.any(p.fvo != 0)
    assert np.allclose(p.fock, p.fock.T)
    pk = p.ovvv_packed()
    assert pk.shape == (3, 9, 45)
    idx, idy = np.tril_indices(9)
    assert np.array_equal(pk, p.ovvv[:, :, idx, idy])


def test_make_L_is_the_factor_behind_make_problem():
    from pyscfad_amd.gen import make_L
    o, v, naux = 3, 7, 11
    L = make_L(o, v, naux, seed=2)
    p = make_problem(o, v, naux, seed=2)
    assert np.array_equal(L, np.swapaxes(L, 1, 2))
    Lov = L[:, :o, o:].reshape(naux, -1)
    assert np.array_equal((Lov.T @ Lov).reshape(o, v, o, v), p.ovov)
