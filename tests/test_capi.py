"""The C-ABI library: it loads, exports every declared symbol, and validates
arguments on the host (no GPU work is started by these calls)."""
import ctypes
import glob
import os
import re

import pytest

from pyscfad_amd import _lib
from tests.conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        text = open(h).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        names.update(re.findall(r'\b(pamd_\w+)\s*\(', text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared()
    assert {'pamd_ccsdt_energy', 'pamd_ccsdt_energy_vjp', 'pamd_ccsdt_workspace_size'} <= declared
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) <= declared


def test_version_and_workspace_size():
    lib = _lib.load()
    assert b'gfx950' in lib.pamd_version()
    cfg = _lib.make_config(8, 40)
    fwd = _lib.workspace_size(cfg, False)
    vjp = _lib.workspace_size(cfg, True)
    assert 0 < fwd < vjp
    # o40/v400: forward holds Vp (o v^3), T2m, OOt and one W block
    big = _lib.workspace_size(_lib.make_config(40, 400), False)
    assert big >= 8 * (40 * 400 ** 3 + 40 * 40 * 400 ** 2 + 400 ** 3)


def test_argument_validation_without_gpu():
    lib = _lib.load()
    null = ctypes.c_void_p(0)
    bad = _lib.make_config(0, 10)
    st = lib.pamd_ccsdt_energy(ctypes.byref(bad), *([null] * 7), null, 0, null, null)
    assert st == _lib.PAMD_ERR_ARG
    cfg = _lib.make_config(4, 10)
    st = lib.pamd_ccsdt_energy(ctypes.byref(cfg), *([null] * 7), null, 0, null, null)
    assert st == _lib.PAMD_ERR_ARG
    with pytest.raises(ValueError):
        _lib.check(st, 'pamd_ccsdt_energy')
    with pytest.raises(ValueError):
        _lib.make_config(4, 10, precision='fp16')


def test_config_struct_layout():
    assert ctypes.sizeof(_lib.Config) == 24


def test_mlp_argument_validation_without_gpu():
    from pyscfad_amd.screen import mlp
    null = ctypes.c_void_p(0)
    fake = ctypes.c_void_p(4096)   # never dereferenced: every call below is rejected first
    gemm = mlp._fn('pamd_bf16_gemm_tn')
    # K not a multiple of 8, lda < K, misaligned operand, missing operand
    assert gemm(4, 4, 12, fake, 16, fake, 16, null, fake, 4, 0, null) == _lib.PAMD_ERR_ARG
    assert gemm(4, 4, 16, fake, 8, fake, 16, null, fake, 4, 0, null) == _lib.PAMD_ERR_ARG
    assert gemm(4, 4, 16, ctypes.c_void_p(4098), 16, fake, 16, null, fake, 4, 0,
                null) == _lib.PAMD_ERR_ARG
    assert gemm(4, 4, 16, null, 16, fake, 16, null, fake, 4, 0, null) == _lib.PAMD_ERR_ARG
    act = mlp._fn('pamd_bf16_act_fwd')
    assert act(4, 4, fake, 2, fake, null, null) == _lib.PAMD_ERR_ARG
    adam = mlp._fn('pamd_adam_step')
    assert adam(2, 2, fake, fake, fake, fake, null, null, 1e-3, 0.9, 0.999, 1e-8, 0.0, 0,
                null) == _lib.PAMD_ERR_ARG
    conv = mlp._fn('pamd_f32_to_bf16')
    assert conv(15, fake, fake, 4, 4, null, null) == _lib.PAMD_ERR_ARG
