"""DF integral producer on the CPU: every contraction descriptor of
pamd_df_eris / pamd_df_eris_vjp evaluated by the emulator (bounds and
write-race checks) and compared with the oracle, before any kernel runs.  The
streaming kernels around the contractions (Lvv packing, zeroing of the
occupied rows, scatter of the Lvv cotangent) are emulated with numpy."""
import numpy as np
import pytest

from oracle import eris_oracle as EO
from tests.gemm_emulator import DF_NAMES, Memory, df_sizes, record_df


def _L(naux, nmo, seed):
    rng = np.random.default_rng(seed)
    G = rng.normal(0, 0.3, (naux, nmo, nmo))
    return np.triu(G) + np.swapaxes(np.triu(G, 1), 1, 2)


CASES = [(3, 9, 5, 0), (3, 9, 5, 1), (4, 10, 6, 3), (1, 5, 3, 0), (5, 7, 4, 2)]


@pytest.mark.parametrize('o,nmo,naux,block', CASES)
def test_df_descriptors_match_the_oracle(o, nmo, naux, block):
    v = nmo - o
    L = _L(naux, nmo, seed=o * 100 + nmo)
    mem = Memory(names=DF_NAMES, array_sizes=df_sizes(o, nmo, naux))
    mem.arrays['Lpq'][:] = L.ravel()
    mem.arrays['Lvv'][:] = EO.pack_tril(L[:, o:, o:]).ravel()        # k_pack_lvv
    for d in record_df(o, nmo, naux, False):
        mem.run(d)
    ref = EO.df_eris(L, o)
    for name, r in zip(('ovvv', 'ovoo', 'ovov'), ref):
        np.testing.assert_allclose(mem.arrays[name].reshape(r.shape), r, rtol=1e-13, atol=1e-14)

    rng = np.random.default_rng(7)
    bars = [rng.normal(size=r.shape) for r in ref]
    for name, b in zip(('ovvv', 'ovoo', 'ovov'), bars):
        mem.arrays[name][:] = b.ravel()
    Lb = mem.arrays['Lpq_bar'].reshape(naux, nmo, nmo)
    Lb[:] = np.nan
    Lb[:, :o, :] = 0.0                                                  # k_zero_occ_rows
    mem.arrays['Lvv_bar'][:] = np.nan
    descs = record_df(o, nmo, naux, True, block_i=block)
    n_blocks = -(-o // (block if block else o))
    assert len(descs) == 1 + 2 * n_blocks
    for d in descs:
        mem.run(d)
    Lvv_bar = mem.arrays['Lvv_bar'].reshape(naux, -1)
    Lb[:, o:, :] = 0.0                                                  # k_unpack_lvv_bar
    Lb[:, o:, o:] = EO.pack_tril_vjp(Lvv_bar, v)
    np.testing.assert_allclose(Lb, EO.df_eris_vjp(L, o, *bars), rtol=1e-12, atol=1e-13)


def test_df_shape_limits_are_checked_on_the_host():
    import ctypes
    from pyscfad_amd import _lib
    lib = _lib.load()
    n = ctypes.c_size_t()
    assert lib.pamd_df_eris_workspace_size(0, 10, 5, 0, ctypes.byref(n)) == _lib.PAMD_ERR_ARG
    assert lib.pamd_df_eris_workspace_size(10, 10, 5, 0, ctypes.byref(n)) == _lib.PAMD_ERR_ARG
    assert lib.pamd_df_eris_workspace_size(40, 440, 1600, 1, ctypes.byref(n)) == 0
    v = 400
    assert n.value >= 2 * 8 * 1600 * v * (v + 1) // 2
    # naux nmo^2 beyond 2^31 elements is refused rather than mis-addressed
    assert lib.pamd_df_eris_workspace_size(40, 2000, 1000, 0, ctypes.byref(n)) == \
        _lib.PAMD_ERR_UNSUPPORTED
    null = ctypes.c_void_p(0)
    assert lib.pamd_df_eris(4, 10, 5, null, null, null, null, null, 0, null) == _lib.PAMD_ERR_ARG
    assert lib.pamd_mp2_init_amps(0, 10, null, null, null, null, null) == _lib.PAMD_ERR_ARG
