"""CPU emulation of the contraction-engine descriptors (test helper).

``record(o, v, i, j, k, with_vjp)`` asks the library (host-only entry point
``pamd_debug_gemm_descriptors``) for the GEMM descriptors of one occupied triple,
with fake base addresses for every workspace array.  ``Memory.run`` then
evaluates each descriptor with numpy gathers, checking that every operand and
output address lies inside its array and that no two outputs of one launch
share an address (no write race).
"""
import ctypes

import numpy as np

from pyscfad_amd import _lib

NAMES = ('Vp', 'T2m', 'OOt', 'W', 't2', 'Vb', 'Vpb', 'T2mb', 'part_p', 'part_h')
BIG = 1 << 30


class OpDesc(ctypes.Structure):
    _fields_ = [('p', ctypes.c_uint64), ('rx', ctypes.c_longlong), ('ry', ctypes.c_longlong),
                ('kx', ctypes.c_longlong), ('ky', ctypes.c_longlong), ('rdiv', ctypes.c_int),
                ('kdiv', ctypes.c_int), ('rfast', ctypes.c_int), ('pad_', ctypes.c_int),
                ('ps', ctypes.c_longlong)]


class SegDesc(ctypes.Structure):
    _fields_ = [('A', OpDesc), ('B', OpDesc), ('L', ctypes.c_int), ('pad_', ctypes.c_int),
                ('alpha', ctypes.c_double)]


class OutDesc(ctypes.Structure):
    _fields_ = [('p', ctypes.c_uint64), ('mx', ctypes.c_longlong), ('my', ctypes.c_longlong),
                ('nx', ctypes.c_longlong), ('ny', ctypes.c_longlong), ('mdiv', ctypes.c_int),
                ('ndiv', ctypes.c_int)]


class GemmDesc(ctypes.Structure):
    _fields_ = [('M', ctypes.c_int), ('N', ctypes.c_int), ('nseg', ctypes.c_int),
                ('mode', ctypes.c_int), ('chunk_start', ctypes.c_int * 5),
                ('splits', ctypes.c_int), ('pad_', ctypes.c_int), ('seg', SegDesc * 4),
                ('C', OutDesc), ('partial', ctypes.c_uint64)]


def sizes(o, v):
    S_p, S_h = split_factors(o, v)
    v2, v3 = v * v, v ** 3
    return dict(Vp=o * v3, T2m=o * o * v2, OOt=o ** 3 * v, W=v3, t2=o * o * v2, Vb=v3,
                Vpb=o * v3, T2mb=o * o * v2, part_p=6 * S_p * v2, part_h=6 * S_h * v * o)


def split_factors(o, v):
    lib = _lib.load()
    lib.pamd_debug_split_factors.argtypes = [ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]
    a, b = ctypes.c_int(), ctypes.c_int()
    assert lib.pamd_debug_split_factors(o, v, ctypes.byref(a), ctypes.byref(b)) == 0
    return a.value, b.value


def record(o, v, i, j, k, with_vjp):
    lib = _lib.load()
    bases = (ctypes.c_ulonglong * 10)(*[(n + 1) << 40 for n in range(10)])
    f = lib.pamd_debug_gemm_descriptors
    f.argtypes = [ctypes.c_int] * 6 + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    n = ctypes.c_int()
    nb = ctypes.c_int()
    buf = (GemmDesc * 64)()
    st = f(o, v, i, j, k, int(with_vjp), bases, buf, 64, ctypes.byref(n), ctypes.byref(nb))
    assert st == 0, st
    assert nb.value == ctypes.sizeof(GemmDesc), (nb.value, ctypes.sizeof(GemmDesc))
    return [buf[q] for q in range(n.value)]


def record_paired(o, v, i, j, kp, k):
    """Adjoint descriptors of the triples (i, j, kp), (i, j, k) as the two-stream
    walk pairs them (pamd_debug_paired_adjoint_descriptors); the second
    triple's Wbar is the array 'W2' (base address after the ten of NAMES)."""
    lib = _lib.load()
    bases = (ctypes.c_ulonglong * 11)(*[(n + 1) << 40 for n in range(11)])
    f = lib.pamd_debug_paired_adjoint_descriptors
    f.argtypes = [ctypes.c_int] * 6 + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    n = ctypes.c_int()
    nb = ctypes.c_int()
    buf = (GemmDesc * 64)()
    st = f(o, v, i, j, kp, k, bases, buf, 64, ctypes.byref(n), ctypes.byref(nb))
    assert st == 0, st
    assert nb.value == ctypes.sizeof(GemmDesc)
    return [buf[q] for q in range(n.value)]


def _decomp(idx, div, sx, sy):
    q = idx // div
    return q * sx + (idx - q * div) * sy


DF_NAMES = ('Lpq', 'Lvv', 'Lvv_bar', 'ovvv', 'ovoo', 'ovov', 'Lpq_bar')


def df_sizes(o, nmo, naux):
    v = nmo - o
    npair = v * (v + 1) // 2
    return dict(Lpq=naux * nmo * nmo, Lvv=naux * npair, Lvv_bar=naux * npair,
                ovvv=o * v * npair, ovoo=o * v * o * o, ovov=(o * v) ** 2,
                Lpq_bar=naux * nmo * nmo)


def record_df(o, nmo, naux, with_vjp, block_i=0):
    """Descriptors of pamd_df_eris (with_vjp = False) or pamd_df_eris_vjp."""
    lib = _lib.load()
    bases = (ctypes.c_ulonglong * 7)(*[(n + 1) << 40 for n in range(7)])
    n = ctypes.c_int()
    nb = ctypes.c_int()
    buf = (GemmDesc * 64)()
    st = lib.pamd_debug_df_descriptors(o, nmo, naux, int(with_vjp), block_i, bases, buf, 64,
                                       ctypes.byref(n), ctypes.byref(nb))
    assert st == 0, st
    assert nb.value == ctypes.sizeof(GemmDesc)
    return [buf[q] for q in range(n.value)]


class Memory:
    def __init__(self, o=None, v=None, names=NAMES, array_sizes=None):
        if array_sizes is None:
            array_sizes = sizes(o, v)
        self.arrays = {name: np.zeros(array_sizes[name]) for name in names}
        self.base = {name: (n + 1) << 40 for n, name in enumerate(names)}

    def locate(self, addr_bytes):
        """map byte addresses to (array name, element offsets); all must hit one array"""
        addr_bytes = np.asarray(addr_bytes, dtype=np.int64)
        lo, hi = int(addr_bytes.min()), int(addr_bytes.max())
        for name, b in self.base.items():
            size = self.arrays[name].size
            if b <= lo < b + 8 * size:
                assert hi < b + 8 * size, f'{name}: address beyond the array'
                off = addr_bytes - b
                assert np.all(off % 8 == 0)
                return name, off // 8
        raise AssertionError(f'address {lo:#x} outside every array')

    def gather(self, p, off):
        name, idx = self.locate(np.int64(p) + 8 * off.astype(np.int64))
        return self.arrays[name][idx]

    def run(self, d):
        M, N = d.M, d.N
        C = np.zeros((M, N))
        m = np.arange(M, dtype=np.int64)
        n = np.arange(N, dtype=np.int64)
        for s in range(d.nseg):
            sg = d.seg[s]
            kk = np.arange(sg.L, dtype=np.int64)
            A, B = sg.A, sg.B
            aoff = _decomp(m, A.rdiv, A.rx, A.ry)[:, None] + _decomp(kk, A.kdiv, A.kx, A.ky)[None, :]
            boff = _decomp(kk, B.kdiv, B.kx, B.ky)[:, None] + _decomp(n, B.rdiv, B.rx, B.ry)[None, :]
            C += sg.alpha * (self.gather(A.p, aoff) @ self.gather(B.p, boff))
        if d.mode == 2:
            name, idx = self.locate(np.int64(d.partial) + 8 * np.arange(d.splits * M * N))
            arr = self.arrays[name]
            arr[idx] = 0.0
            arr[idx[:M * N]] = C.reshape(-1)
        else:
            coff = (_decomp(m, d.C.mdiv, d.C.mx, d.C.my)[:, None]
                    + _decomp(n, d.C.ndiv, d.C.nx, d.C.ny)[None, :])
            name, idx = self.locate(np.int64(d.C.p) + 8 * coff)
            flat = idx.reshape(-1)
            assert np.unique(flat).size == flat.size, 'two outputs of one launch share an address'
            arr = self.arrays[name]
            if d.mode == 1:
                arr[flat] += C.reshape(-1)
            else:
                arr[flat] = C.reshape(-1)
        return d
