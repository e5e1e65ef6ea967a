"""Build libpyscfad_amd_cc.so for gfx950 with hipcc (in-tree).

    python -m pyscfad_amd.csrc.build

Every HIP translation unit is compiled with packed-FP32 code generation turned
off (``-Xclang -target-feature -Xclang -packed-fp32-ops``): no v_pk_*_f32 is
emitted anywhere in the library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB_NAME = 'libpyscfad_amd_cc.so'
LIB_PATH = os.path.join(PKG, LIB_NAME)
SOURCES = ('ccsdt_kernels.hip', 'gemm_direct.hip', 'gemm_b3p.hip', 'bf16_kernels.hip', 'mlp_kernels.hip', 'eris_kernels.hip',
           'ccsdt_driver.cpp', 'mlp_capi.cpp', 'eris_driver.cpp')
ARCH = os.environ.get('PYSCFAD_AMD_ARCH', 'gfx950')


def hipcc():
    for cand in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose: bool = False) -> str:
    objdir = os.path.join(ROOT, '_build')
    os.makedirs(objdir, exist_ok=True)
    common = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
              '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
              '-I', os.path.join(ROOT, 'include'), '-I', HERE]
    objs = []
    for src in SOURCES:
        path = os.path.join(HERE, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + '.o')
        lang = ['-x', 'hip'] if src.endswith('.cpp') else []
        cmd = [hipcc()] + lang + common + ['-c', path, '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f'hipcc failed on {src}')
        objs.append(obj)
    tmp = LIB_PATH + '.tmp'
    cmd = [hipcc(), '-shared', f'--offload-arch={ARCH}', '-o', tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('link failed')
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    print(build(verbose=True))
