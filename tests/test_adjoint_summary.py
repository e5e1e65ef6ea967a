"""The error measures of the finite-difference adjoint check, on values whose
answers are known (no GPU)."""
import math

import pytest


def test_summarise_known_values():
    from pyscfad_amd.cc.adjoint_check import summarise
    vals = {'fd': [1.0, -1.0, 2.0], 'fp64': [1.0, -1.0, 2.0], 'bf16x3': [1.1, -1.0, 2.0]}
    s = summarise(vals, ['fp64', 'bf16x3'])
    rms = math.sqrt(2.0)
    assert s['fp64']['rms_derivative'] == pytest.approx(rms)
    assert s['fp64']['err_vs_rms_max'] == 0.0
    assert 'vs_fp64_rms_max' not in s['fp64']
    assert s['bf16x3']['err_vs_rms_max'] == pytest.approx(0.1 / rms)
    assert s['bf16x3']['err_vs_rms_median'] == 0.0
    assert s['bf16x3']['vs_fp64_rms_max'] == pytest.approx(0.1 / rms)


def test_summarise_single_precision_uses_its_own_rms():
    from pyscfad_amd.cc.adjoint_check import summarise
    s = summarise({'fd': [3.0, 4.0], 'bf16x3': [3.0, 4.5]}, ['bf16x3'])
    assert s['bf16x3']['rms_derivative'] == pytest.approx(math.sqrt((9.0 + 20.25) / 2))
    assert 'vs_fp64_rms_max' not in s['bf16x3']
