"""Adjoint of both precisions against central finite differences of the FP64
GPU energy, along seeded unit directions in the joint input space."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
def test_central_fd_directions(cuda, precision):
    from pyscfad_amd.cc import adjoint_check
    rep = adjoint_check.run(6, 32, 128, directions=8, h=1e-5, precisions=[precision],
                            log=lambda s: print(s, flush=True))
    r = rep['precisions'][precision]
    print(f"{precision}: max rel err {r['rel_err_max']:.2e}, median {r['rel_err_median']:.2e}")
    # FP64: FD truncation/rounding only; split-BF16x3: its gradient error (<~1e-7)
    assert r['rel_err_max'] <= (1e-6 if precision == 'fp64' else 1e-5)
    # the VJP keeps one v^3 block of W/V, not all triples
    assert r['peak_device_bytes_vjp'] < r['store_everything_bytes']
