"""Device-level (T) operations on torch tensors, calling the C ABI.

These functions take device (or host) arrays in the reference's layouts, move
them to the GPU as contiguous float64 if needed, and run the HIP kernels on the
current torch stream.  There is no CPU path: without a GPU they raise.
"""
from __future__ import annotations

import ctypes
import numpy as np
import torch

from .. import _lib


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('pyscfad_amd: no GPU is visible; the (T) kernels are HIP-only '
                           'and have no CPU path')


def _dev(x, device):
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        t = x.detach()
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64)))
    return t.to(device=device, dtype=torch.float64).contiguous()


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Inputs:
    """Device copies of the (T) inputs plus the C-ABI config."""

    def __init__(self, mo_energy, t1, t2, ovvv, ovoo, ovov, fvo=None, precision='fp64',
                 device=None):
        _require_gpu()
        if device is None:
            device = t2.device if isinstance(t2, torch.Tensor) and t2.is_cuda else \
                torch.device('cuda', torch.cuda.current_device())
        self.device = torch.device(device)
        self.t1 = _dev(t1, self.device)
        nocc, nvir = self.t1.shape
        self.nocc, self.nvir = int(nocc), int(nvir)
        self.mo_energy = _dev(mo_energy, self.device)
        self.t2 = _dev(t2, self.device)
        self.ovvv = _dev(ovvv, self.device)
        self.ovoo = _dev(ovoo, self.device)
        self.ovov = _dev(ovov, self.device)
        self.fvo = _dev(fvo, self.device)
        o, v = self.nocc, self.nvir
        self.packed = self.ovvv.dim() == 3
        expect = {
            'mo_energy': (self.mo_energy, (o + v,)),
            't2': (self.t2, (o, o, v, v)),
            'ovvv': (self.ovvv, (o, v, v * (v + 1) // 2) if self.packed else (o, v, v, v)),
            'ovoo': (self.ovoo, (o, v, o, o)),
            'ovov': (self.ovov, (o, v, o, v)),
        }
        if self.fvo is not None:
            expect['fvo'] = (self.fvo, (v, o))
        for name, (t, shape) in expect.items():
            if tuple(t.shape) != shape:
                raise ValueError(f'{name} has shape {tuple(t.shape)}, expected {shape}')
        self.has_fvo = self.fvo is not None and bool(torch.count_nonzero(self.fvo).item() > 0)
        if not self.has_fvo:
            self.fvo = None
        self.precision = precision
        self.cfg = _lib.make_config(o, v, self.packed, self.has_fvo, precision)

    def args(self):
        return [_ptr(self.mo_energy), _ptr(self.t1), _ptr(self.t2), _ptr(self.ovvv),
                _ptr(self.ovoo), _ptr(self.ovov), _ptr(self.fvo)]

    def workspace(self, with_vjp):
        # no device query before the allocation: a query can wait for the
        # device, which would serialise calls issued on several streams
        nbytes = _lib.workspace_size(self.cfg, with_vjp)
        try:
            ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)
        except torch.OutOfMemoryError as exc:
            raise RuntimeError(f'(T): {nbytes / 1e6:.0f} MB of device workspace needed') from exc
        return ws, nbytes


def energy(inp: Inputs):
    """E(T) as a 0-d float64 device tensor."""
    lib = _lib.load()
    ws, nbytes = inp.workspace(False)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    st = lib.pamd_ccsdt_energy(ctypes.byref(inp.cfg), *inp.args(), _ptr(ws), nbytes, _ptr(e),
                               _stream(inp.device))
    _lib.check(st, 'pamd_ccsdt_energy')
    return e


def energy_shard(inp: Inputs, shard, nshards):
    """E(T) of the round-robin shard ``shard`` of ``nshards`` over the occupied
    triples (the shards add up to ``energy``)."""
    lib = _lib.load()
    ws, nbytes = inp.workspace(False)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    st = lib.pamd_ccsdt_energy_shard(ctypes.byref(inp.cfg), *inp.args(), int(shard),
                                     int(nshards), _ptr(ws), nbytes, _ptr(e),
                                     _stream(inp.device))
    _lib.check(st, 'pamd_ccsdt_energy_shard')
    return e


def pair_energy(inp: Inputs):
    """(o, o) occupied-pair resolved E(T); its sum is E(T)."""
    lib = _lib.load()
    ws, nbytes = inp.workspace(False)
    e = torch.zeros((inp.nocc, inp.nocc), dtype=torch.float64, device=inp.device)
    st = lib.pamd_ccsdt_pair_energy(ctypes.byref(inp.cfg), *inp.args(), _ptr(ws), nbytes,
                                    _ptr(e), _stream(inp.device))
    _lib.check(st, 'pamd_ccsdt_pair_energy')
    return e


def _cotangent_outputs(inp: Inputs):
    o, v = inp.nocc, inp.nvir
    f64 = dict(dtype=torch.float64, device=inp.device)
    return dict(
        mo_energy_bar=torch.empty(o + v, **f64),
        t1_bar=torch.empty((o, v), **f64),
        t2_bar=torch.empty((o, o, v, v), **f64),
        ovvv_bar=torch.empty(tuple(inp.ovvv.shape), **f64),
        ovoo_bar=torch.empty((o, v, o, o), **f64),
        ovov_bar=torch.empty((o, v, o, v), **f64),
        fvo_bar=torch.empty((v, o), **f64),
    )


def _cotangent_ptrs(out):
    return [_ptr(out[k]) for k in ('mo_energy_bar', 't1_bar', 't2_bar', 'ovvv_bar', 'ovoo_bar',
                                   'ovov_bar', 'fvo_bar')]


def energy_vjp(inp: Inputs, et_bar=1.0, workspace=None):
    """(E(T), cotangents) with the cotangents in the reference's output layout.
    ``workspace``: an optional (tensor, nbytes) from ``inp.workspace(True)``,
    reused across calls."""
    lib = _lib.load()
    ws, nbytes = workspace if workspace is not None else inp.workspace(True)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    out = _cotangent_outputs(inp)
    st = lib.pamd_ccsdt_energy_vjp(
        ctypes.byref(inp.cfg), *inp.args(), ctypes.c_double(float(et_bar)), _ptr(ws), nbytes,
        _ptr(e), *_cotangent_ptrs(out), _stream(inp.device))
    _lib.check(st, 'pamd_ccsdt_energy_vjp')
    return e, out


def energy_vjp_shard(inp: Inputs, shard, nshards, et_bar=1.0, workspace=None):
    """(E, cotangents) of the round-robin shard ``shard`` of ``nshards`` over the
    occupied triples; the shards' energies and cotangents add up to
    ``energy_vjp``'s."""
    lib = _lib.load()
    ws, nbytes = workspace if workspace is not None else inp.workspace(True)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    out = _cotangent_outputs(inp)
    st = lib.pamd_ccsdt_energy_vjp_shard(
        ctypes.byref(inp.cfg), *inp.args(), ctypes.c_double(float(et_bar)), int(shard),
        int(nshards), _ptr(ws), nbytes, _ptr(e), *_cotangent_ptrs(out), _stream(inp.device))
    _lib.check(st, 'pamd_ccsdt_energy_vjp_shard')
    return e, out


def weighted_energy_vjp(inp: Inputs, occ_weight, et_bar=1.0):
    """(E_m, cotangents) for the diagonal occupied projector diag(m):
    E_m = sum_i m_i E_i with E_i the row sums of ``pair_energy``; the cotangents
    include ``occ_weight_bar`` = et_bar * E_i."""
    lib = _lib.load()
    o, v = inp.nocc, inp.nvir
    m = np.ascontiguousarray(np.asarray(
        occ_weight.detach().cpu() if isinstance(occ_weight, torch.Tensor) else occ_weight,
        dtype=np.float64))
    if m.shape != (o,):
        raise ValueError(f'occ_weight must have {o} entries, got {m.shape}')
    dev = inp.device
    ws, nbytes = inp.workspace(True)
    f64 = dict(dtype=torch.float64, device=dev)
    e = torch.zeros((), **f64)
    out = dict(
        mo_energy_bar=torch.empty(o + v, **f64),
        t1_bar=torch.empty((o, v), **f64),
        t2_bar=torch.empty((o, o, v, v), **f64),
        ovvv_bar=torch.empty(tuple(inp.ovvv.shape), **f64),
        ovoo_bar=torch.empty((o, v, o, o), **f64),
        ovov_bar=torch.empty((o, v, o, v), **f64),
        fvo_bar=torch.empty((v, o), **f64),
        occ_weight_bar=torch.empty(o, **f64),
    )
    st = lib.pamd_ccsdt_weighted_energy_vjp(
        ctypes.byref(inp.cfg), *inp.args(), m.ctypes.data_as(ctypes.c_void_p),
        ctypes.c_double(float(et_bar)), _ptr(ws), nbytes, _ptr(e), _ptr(out['mo_energy_bar']),
        _ptr(out['t1_bar']), _ptr(out['t2_bar']), _ptr(out['ovvv_bar']), _ptr(out['ovoo_bar']),
        _ptr(out['ovov_bar']), _ptr(out['fvo_bar']), _ptr(out['occ_weight_bar']), _stream(dev))
    _lib.check(st, 'pamd_ccsdt_weighted_energy_vjp')
    torch.cuda.current_stream(dev).synchronize()   # m is read from host memory
    return e, out


def _ulo_array(inp, ulo):
    u = np.ascontiguousarray(np.asarray(
        ulo.detach().cpu() if isinstance(ulo, torch.Tensor) else ulo, dtype=np.float64))
    if u.ndim == 1:
        u = u[None, :]
    if u.ndim != 2 or u.shape[1] != inp.nocc:
        raise ValueError(f'ulo must be (nfrag, {inp.nocc}), got {u.shape}')
    return u


def _projected_ws(inp, nfrag, with_vjp):
    lib = _lib.load()
    n = ctypes.c_size_t(0)
    _lib.check(lib.pamd_ccsdt_projected_workspace_size(ctypes.byref(inp.cfg), nfrag,
                                                       int(with_vjp), ctypes.byref(n)),
               'pamd_ccsdt_projected_workspace_size')
    return torch.empty(max(n.value, 1), dtype=torch.uint8, device=inp.device), n.value


# fragment orbitals per library call (nfrag <= 8 in pamd_ccsdt_projected_*); E is
# linear in mat = ulo^T ulo = sum_f u_f u_f^T, so larger fragments are summed
# over row groups of ulo
MAX_FRAG_PER_CALL = 8


def _row_groups(u):
    return [u[r:r + MAX_FRAG_PER_CALL] for r in range(0, u.shape[0], MAX_FRAG_PER_CALL)]


def projected_energy(inp: Inputs, ulo):
    """Occupied-weighted E(T) of the local-orbital projector mat = ulo^T ulo
    (ulo: (nfrag, nocc), any nfrag >= 1), as a 0-d float64 device tensor."""
    lib = _lib.load()
    u = _ulo_array(inp, ulo)
    e = torch.zeros((), dtype=torch.float64, device=inp.device)
    for ug in _row_groups(u):
        ug = np.ascontiguousarray(ug)
        ws, nb = _projected_ws(inp, ug.shape[0], False)
        eg = torch.zeros((), dtype=torch.float64, device=inp.device)
        st = lib.pamd_ccsdt_projected_energy(ctypes.byref(inp.cfg), *inp.args(),
                                             ug.ctypes.data_as(ctypes.c_void_p), int(ug.shape[0]),
                                             _ptr(ws), nb, _ptr(eg), _stream(inp.device))
        _lib.check(st, 'pamd_ccsdt_projected_energy')
        torch.cuda.current_stream(inp.device).synchronize()   # ug is read from host memory
        e += eg
    return e


def projected_energy_vjp(inp: Inputs, ulo, et_bar=1.0):
    """(E, cotangents) of ``projected_energy``; the cotangents include
    ``ulo_bar`` (nfrag, nocc).  Fragments of more than MAX_FRAG_PER_CALL orbitals
    run as row groups of ulo: the energies and the shared cotangents add, and
    the rows of ulo_bar = ulo (mat_bar + mat_bar^T) come from each group's own
    rows (mat_bar does not depend on mat)."""
    lib = _lib.load()
    o, v = inp.nocc, inp.nvir
    u = _ulo_array(inp, ulo)
    dev = inp.device
    f64 = dict(dtype=torch.float64, device=dev)
    e = torch.zeros((), **f64)
    out = None
    ulo_bar = []
    for ug in _row_groups(u):
        ug = np.ascontiguousarray(ug)
        ws, nb = _projected_ws(inp, ug.shape[0], True)
        eg = torch.zeros((), **f64)
        g = dict(
            mo_energy_bar=torch.empty(o + v, **f64),
            t1_bar=torch.empty((o, v), **f64),
            t2_bar=torch.empty((o, o, v, v), **f64),
            ovvv_bar=torch.empty(tuple(inp.ovvv.shape), **f64),
            ovoo_bar=torch.empty((o, v, o, o), **f64),
            ovov_bar=torch.empty((o, v, o, v), **f64),
            fvo_bar=torch.empty((v, o), **f64),
            ulo_bar=torch.empty(tuple(ug.shape), **f64),
        )
        st = lib.pamd_ccsdt_projected_energy_vjp(
            ctypes.byref(inp.cfg), *inp.args(), ug.ctypes.data_as(ctypes.c_void_p),
            int(ug.shape[0]), ctypes.c_double(float(et_bar)), _ptr(ws), nb, _ptr(eg),
            _ptr(g['mo_energy_bar']), _ptr(g['t1_bar']), _ptr(g['t2_bar']), _ptr(g['ovvv_bar']),
            _ptr(g['ovoo_bar']), _ptr(g['ovov_bar']), _ptr(g['fvo_bar']), _ptr(g['ulo_bar']),
            _stream(dev))
        _lib.check(st, 'pamd_ccsdt_projected_energy_vjp')
        torch.cuda.current_stream(dev).synchronize()   # ug is read from host memory
        del ws
        e += eg
        ulo_bar.append(g.pop('ulo_bar'))
        if out is None:
            out = g
        else:
            for k in out:
                out[k] += g[k]
    out['ulo_bar'] = torch.cat(ulo_bar, 0)
    return e, out
