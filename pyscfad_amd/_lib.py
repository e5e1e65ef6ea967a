"""ctypes binding of libpyscfad_amd_cc.so (the C ABI in include/pyscfad_amd_ccsdt.h).

The library is built in-tree by ``python -m pyscfad_amd.csrc.build``.  There is
no CPU fallback: every product entry point raises if the library or a GPU is
missing.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libpyscfad_amd_cc.so')

PAMD_OK = 0
PAMD_ERR_ARG = 1
PAMD_ERR_HIP = 2
PAMD_ERR_WORKSPACE = 3
PAMD_ERR_UNSUPPORTED = 4
STATUS_NAMES = {0: 'ok', 1: 'bad argument', 2: 'HIP runtime error', 3: 'workspace too small',
                4: 'unsupported option'}

PRECISIONS = {'fp64': 0, 'bf16x3': 1, 'fp32x2': 2}

# exported symbols of include/pyscfad_amd_ccsdt.h, _profile.h, _debug.h, _mlp.h and _eris.h
EXPORTS = ('pamd_version', 'pamd_ccsdt_workspace_size', 'pamd_ccsdt_energy',
           'pamd_ccsdt_energy_vjp', 'pamd_ccsdt_pair_energy', 'pamd_ccsdt_num_triples',
           'pamd_ccsdt_energy_shard', 'pamd_ccsdt_energy_vjp_shard',
           'pamd_ccsdt_energy_vjp_noalloc', 'pamd_ccsdt_noalloc_scratch_bytes',
           'pamd_ccsdt_weighted_energy_vjp',
           'pamd_ccsdt_projected_workspace_size', 'pamd_ccsdt_projected_energy',
           'pamd_ccsdt_projected_energy_vjp', 'pamd_profile_enable',
           'pamd_profile_reset', 'pamd_profile_progress', 'pamd_profile_num_kinds', 'pamd_profile_kind_name',
           'pamd_profile_read', 'pamd_debug_gemm_descriptors',
           'pamd_debug_paired_adjoint_descriptors', 'pamd_debug_split_factors',
           'pamd_debug_split_factors_prec', 'pamd_debug_gemm_direct_ok', 'pamd_debug_gemm_b3p_ok',
           'pamd_bf16_gemm_tn', 'pamd_bf16_act_fwd', 'pamd_bf16_act_bwd', 'pamd_adam_step',
           'pamd_mse', 'pamd_f32_to_bf16', 'pamd_mp2_init_amps', 'pamd_df_eris_workspace_size',
           'pamd_df_eris', 'pamd_df_eris_vjp', 'pamd_debug_df_descriptors')


class Config(ctypes.Structure):
    _fields_ = [('nocc', ctypes.c_int), ('nvir', ctypes.c_int), ('ovvv_packed', ctypes.c_int),
                ('has_fvo', ctypes.c_int), ('precision', ctypes.c_int),
                ('reserved', ctypes.c_int)]


_lib = None


def load():
    """Load the HIP library (raises RuntimeError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'{LIB_PATH} is missing: build it with '
                           '`python -m pyscfad_amd.csrc.build` (hipcc, gfx950)')
    lib = ctypes.CDLL(LIB_PATH)
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    cfgp = ctypes.POINTER(Config)
    lib.pamd_version.restype = ctypes.c_char_p
    lib.pamd_version.argtypes = []
    lib.pamd_ccsdt_workspace_size.restype = i
    lib.pamd_ccsdt_workspace_size.argtypes = [cfgp, i, ctypes.POINTER(sz)]
    lib.pamd_ccsdt_energy.restype = i
    lib.pamd_ccsdt_energy.argtypes = [cfgp] + [vp] * 7 + [vp, sz, vp, vp]
    lib.pamd_ccsdt_num_triples.restype = i
    lib.pamd_ccsdt_num_triples.argtypes = [i, ctypes.POINTER(ctypes.c_longlong)]
    lib.pamd_ccsdt_energy_shard.restype = i
    lib.pamd_ccsdt_energy_shard.argtypes = [cfgp] + [vp] * 7 + [i, i, vp, sz, vp, vp]
    lib.pamd_ccsdt_pair_energy.restype = i
    lib.pamd_ccsdt_pair_energy.argtypes = [cfgp] + [vp] * 7 + [vp, sz, vp, vp]
    lib.pamd_ccsdt_energy_vjp.restype = i
    lib.pamd_ccsdt_energy_vjp.argtypes = [cfgp] + [vp] * 7 + [d, vp, sz] + [vp] * 8 + [vp]
    lib.pamd_ccsdt_energy_vjp_shard.restype = i
    lib.pamd_ccsdt_energy_vjp_shard.argtypes = ([cfgp] + [vp] * 7 + [d, i, i, vp, sz] +
                                                [vp] * 8 + [vp])
    lib.pamd_ccsdt_energy_vjp_noalloc.restype = i
    lib.pamd_ccsdt_energy_vjp_noalloc.argtypes = [cfgp] + [vp] * 7 + [d] + [vp] * 8 + [vp]
    lib.pamd_ccsdt_noalloc_scratch_bytes.restype = sz
    lib.pamd_ccsdt_noalloc_scratch_bytes.argtypes = []
    lib.pamd_ccsdt_projected_workspace_size.restype = i
    lib.pamd_ccsdt_projected_workspace_size.argtypes = [cfgp, i, i, ctypes.POINTER(sz)]
    lib.pamd_ccsdt_projected_energy_vjp.restype = i
    lib.pamd_ccsdt_projected_energy_vjp.argtypes = ([cfgp] + [vp] * 8 + [i, d, vp, sz] +
                                                    [vp] * 9 + [vp])
    lib.pamd_ccsdt_projected_energy.restype = i
    lib.pamd_ccsdt_projected_energy.argtypes = [cfgp] + [vp] * 8 + [i, vp, sz, vp, vp]
    lib.pamd_ccsdt_weighted_energy_vjp.restype = i
    lib.pamd_ccsdt_weighted_energy_vjp.argtypes = [cfgp] + [vp] * 8 + [d, vp, sz] + [vp] * 9 + [vp]
    lib.pamd_profile_enable.argtypes = [i]
    lib.pamd_profile_enable.restype = i
    lib.pamd_profile_progress.argtypes = [i]
    lib.pamd_profile_progress.restype = i
    lib.pamd_profile_reset.argtypes = []
    lib.pamd_profile_reset.restype = i
    lib.pamd_profile_num_kinds.argtypes = []
    lib.pamd_profile_num_kinds.restype = i
    lib.pamd_profile_kind_name.argtypes = [i]
    lib.pamd_profile_kind_name.restype = ctypes.c_char_p
    lib.pamd_profile_read.argtypes = [i, ctypes.POINTER(ctypes.c_longlong)] + \
        [ctypes.POINTER(ctypes.c_double)] * 3
    lib.pamd_profile_read.restype = i
    lib.pamd_mp2_init_amps.restype = i
    lib.pamd_mp2_init_amps.argtypes = [i, i] + [vp] * 5
    lib.pamd_df_eris_workspace_size.restype = i
    lib.pamd_df_eris_workspace_size.argtypes = [i, i, i, i, ctypes.POINTER(sz)]
    lib.pamd_df_eris.restype = i
    lib.pamd_df_eris.argtypes = [i, i, i] + [vp] * 5 + [sz, vp]
    lib.pamd_df_eris_vjp.restype = i
    lib.pamd_df_eris_vjp.argtypes = [i, i, i] + [vp] * 6 + [sz, vp]
    lib.pamd_debug_gemm_direct_ok.restype = i
    lib.pamd_debug_gemm_direct_ok.argtypes = [vp]
    lib.pamd_debug_gemm_b3p_ok.restype = i
    lib.pamd_debug_gemm_b3p_ok.argtypes = [vp, ctypes.c_longlong]
    lib.pamd_debug_df_descriptors.restype = i
    lib.pamd_debug_df_descriptors.argtypes = [i, i, i, i, i, vp, vp, i, ctypes.POINTER(i),
                                              ctypes.POINTER(i)]
    _lib = lib
    return lib


def check(status, what):
    if status != PAMD_OK:
        name = STATUS_NAMES.get(status, f'status {status}')
        if status in (PAMD_ERR_ARG, PAMD_ERR_UNSUPPORTED):
            raise ValueError(f'{what}: {name}')
        raise RuntimeError(f'{what}: {name}')


def make_config(nocc, nvir, ovvv_packed=False, has_fvo=False, precision='fp64'):
    if precision not in PRECISIONS:
        raise ValueError(f'unknown precision {precision!r}; choose from {sorted(PRECISIONS)}')
    return Config(int(nocc), int(nvir), int(bool(ovvv_packed)), int(bool(has_fvo)),
                  PRECISIONS[precision], 0)


def workspace_size(cfg, with_vjp):
    lib = load()
    out = ctypes.c_size_t(0)
    check(lib.pamd_ccsdt_workspace_size(ctypes.byref(cfg), int(bool(with_vjp)), ctypes.byref(out)),
          'pamd_ccsdt_workspace_size')
    return out.value


def profile_enable(every):
    """Time the launches of one occupied triple in ``every`` (0 disables)."""
    check(load().pamd_profile_enable(int(every)), 'pamd_profile_enable')


def profile_progress(every):
    """Print a progress line (stderr) every ``every`` occupied triples (0 disables)."""
    check(load().pamd_profile_progress(int(every)), 'pamd_profile_progress')


def profile_reset():
    check(load().pamd_profile_reset(), 'pamd_profile_reset')


def profile_read():
    """{kernel: dict(launches, ms, flops, bytes)} accumulated since the last reset."""
    lib = load()
    out = {}
    for kind in range(lib.pamd_profile_num_kinds()):
        n = ctypes.c_longlong()
        ms, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib.pamd_profile_read(kind, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl),
                                    ctypes.byref(by)), 'pamd_profile_read')
        out[lib.pamd_profile_kind_name(kind).decode()] = dict(
            launches=n.value, ms=ms.value, flops=fl.value, bytes=by.value)
    return out
