"""Learned screen: BF16 MLP kernels against the CPU autograd reference, training,
the spread-gated fallback, and what the trained ensemble gains over running the
exact (T) on every instance.
"""
import time

import numpy as np
import pytest
import torch

from tests import mlp_reference as REF

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(torch.linalg.norm((a - b).double()) / torch.linalg.norm(b.double()))


@pytest.fixture(scope='module')
def M():
    from pyscfad_amd.screen import mlp
    return mlp


def test_gemm_tn_shapes(cuda, M):
    g = torch.Generator().manual_seed(0)
    for (m, n, k) in [(1, 1, 8), (37, 19, 40), (256, 136, 256), (136, 256, 256), (128, 80, 136)]:
        A = REF.bf16_round(torch.randn(m, k, generator=g))
        B = REF.bf16_round(torch.randn(n, k, generator=g))
        bias = torch.randn(n, generator=g)
        A16 = M.to_bf16(A.to(cuda))
        B16 = M.to_bf16(B.to(cuda))
        C = M.gemm_tn(A16, B16, bias=bias.to(cuda)).cpu()
        ref = A.double() @ B.double().T + bias.double()
        assert float((C.double() - ref).abs().max()) <= 1e-5 * (1 + float(ref.abs().max())), (m, n, k)
        # transposed copy written by the conversion kernel
        _, At = M.to_bf16(A.to(cuda), with_t=True)
        assert torch.equal(REF.bits_to_f32(At.cpu()), A.T.contiguous())


def test_bf16_conversion_is_rne(cuda, M):
    x = torch.randn(4096, generator=torch.Generator().manual_seed(1)) * 3.0
    x[:4] = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0])  # ties
    b = M.to_bf16(x.view(64, 64).to(cuda)).cpu()
    assert torch.equal(REF.bits_to_f32(b).flatten(), REF.bf16_round(x))


def _setup(cuda, M, sizes, batch, seed):
    net = M.MLP(sizes, cuda, seed=seed)
    g = torch.Generator().manual_seed(seed + 1)
    x = REF.bf16_round(torch.randn(batch, sizes[0], generator=g))
    y = torch.randn(batch, sizes[-1], generator=g)
    for L in net.layers:  # non-zero biases so that their gradients matter
        L['b'].copy_(0.1 * torch.randn(L['b'].shape, generator=g).to(cuda))
    x16, x16t = M.to_bf16(x.to(cuda), with_t=True)
    return net, x, y, x16, x16t


@pytest.mark.parametrize('sizes,batch', [((128, 256, 256, 136), 256), ((40, 64, 24), 72)])
def test_mlp_forward_backward_vs_cpu_autograd(cuda, M, sizes, batch):
    net, x, y, x16, x16t = _setup(cuda, M, sizes, batch, seed=3)
    pred, loss, grads = net.grads(x16, x16t, y.to(cuda))
    Ws = [REF.bits_to_f32(L['wb'].cpu()) for L in net.layers]
    bs = [L['b'].cpu() for L in net.layers]
    pred_r, loss_r, grads_r = REF.loss_and_grads(Ws, bs, x, y)
    # forward: same bf16 operands and roundings, FP32 sums in another order
    assert _rel(pred.cpu(), pred_r) <= 1e-4
    assert abs(float(loss) - float(loss_r)) <= 1e-4 * float(loss_r)
    # backward: the GPU rounds dZ to bf16 before the weight/input products
    for (dW, db), (dW_r, db_r) in zip(grads, grads_r):
        assert _rel(dW.cpu(), dW_r) <= 1e-2
        assert _rel(db.cpu(), db_r) <= 1e-2


def test_adam_step_vs_reference(cuda, M):
    net, x, y, x16, x16t = _setup(cuda, M, (40, 64, 24), 72, seed=5)
    before = [(L['w'].cpu().clone(), L['b'].cpu().clone()) for L in net.layers]
    _, _, grads = net.grads(x16, x16t, y.to(cuda))
    net.adam(grads, lr=1e-2, wd=1e-3)
    for L, (w0, b0), (dW, db) in zip(net.layers, before, grads):
        w_r, m_r, v_r = REF.adam(w0, dW.cpu(), torch.zeros_like(w0), torch.zeros_like(w0), 1, 1e-2,
                                 wd=1e-3)
        assert float((L['w'].cpu().double() - w_r).abs().max()) <= 1e-6
        assert float((L['m_w'].cpu().double() - m_r).abs().max()) <= 1e-7 * (1 + float(m_r.abs().max()))
        b_r, _, _ = REF.adam(b0, db.cpu(), torch.zeros_like(b0), torch.zeros_like(b0), 1, 1e-2)
        assert float((L['b'].cpu().double() - b_r).abs().max()) <= 1e-6
        # the bf16 copies follow the master weights, row-major and transposed
        wf = REF.bf16_round(L['w'].cpu())
        assert torch.equal(REF.bits_to_f32(L['wb'].cpu()), wf)
        assert torch.equal(REF.bits_to_f32(L['wt'].cpu()), wf.T.contiguous())


def test_training_reduces_loss(cuda, M):
    """A smooth 40 -> 24 regression: the kernels' training lowers the loss 10x."""
    g = torch.Generator().manual_seed(11)
    X = torch.randn(512, 40, generator=g)
    A = torch.randn(40, 24, generator=g) / 40 ** 0.5
    Y = torch.tanh(X @ A) + 0.1 * (X[:, :24] ** 2)
    net = M.MLP((40, 128, 128, 24), cuda, seed=2)
    x16, x16t = M.to_bf16(X.to(cuda), with_t=True)
    Yd = Y.to(cuda)
    first = float(net.mse(net.forward(x16), Yd, want_grad=False)[0])
    for it in range(300):
        sl = slice((it % 4) * 128, (it % 4 + 1) * 128)
        xb = x16[sl].contiguous()
        net.train_step(xb, xb.T.contiguous(), Yd[sl].contiguous(), lr=3e-3)
    last = float(net.mse(net.forward(x16), Yd, want_grad=False)[0])
    assert last < 0.1 * first, (first, last)


def test_fallback_branches_forced(cuda):
    from pyscfad_amd.screen.ensemble import Screen
    g = torch.Generator().manual_seed(4)
    X = torch.randn(64, 12, generator=g, dtype=torch.float64).to(cuda)
    Y = 1e-4 * torch.randn(64, 10, generator=g, dtype=torch.float64).to(cuda)
    sc = Screen(12, 10, cuda, members=2, hidden=(32, 32))
    sc.fit(X, Y, epochs=2, batch=32)
    calls = []

    def exact(i):
        calls.append(i)
        return float(Y[i].sum())

    mean, _ = sc.predict(X)
    e, _, fb, _ = sc.energy(X, exact, spread_override=torch.zeros(64))
    assert fb == [] and calls == [] and torch.equal(e, mean)
    e, _, fb, _ = sc.energy(X, exact, spread_override=torch.ones(64))
    assert fb == list(range(64)) and calls == list(range(64))
    assert torch.allclose(e, Y.sum(1))
    half = torch.zeros(64)
    half[::2] = 1.0
    calls.clear()
    e, _, fb, _ = sc.energy(X, exact, spread_override=half)
    assert fb == list(range(0, 64, 2))
    assert torch.allclose(e[::2], Y.sum(1)[::2]) and torch.equal(e[1::2], mean[1::2])


def test_screen_gain_small_instances(cuda):
    """Perturbed o12/v96 instances: the trained ensemble answers most test
    instances without the exact kernel, and its accepted answers are far more
    accurate than the MP2-scaling estimate available without it."""
    from pyscfad_amd.screen import data as D
    from pyscfad_amd.screen.ensemble import Screen
    from pyscfad_amd.screen.features import features, fold_pairs, n_features, n_labels
    o, v, naux, n, n_train = 12, 96, 384, 448, 384
    iset = D.InstanceSet(cuda, o, v, naux)
    X, Y, E = [], [], []
    t_exact = 0.0
    for s in range(1, n + 1):
        p = iset.instance(s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ep = D.exact_pairs(p)
        torch.cuda.synchronize()
        t_exact += time.perf_counter() - t0
        X.append(features(p.mo_energy, p.ovov, o))
        Y.append(fold_pairs(ep))
        E.append(ep.sum())
    X, Y, E = torch.stack(X), torch.stack(Y), torch.stack(E)
    assert X.shape == (n, n_features(o)) and Y.shape == (n, n_labels(o))
    assert torch.allclose(Y.sum(1), E, rtol=0, atol=1e-15)
    tr, te = D.split(n, n_train)
    tr, te = torch.as_tensor(tr, device=cuda), torch.as_tensor(te, device=cuda)
    sc = Screen(X.shape[1], Y.shape[1], cuda, members=4)
    sc.fit(X[tr], Y[tr], epochs=150, batch=64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mean, spread = sc.predict(X[te])
    torch.cuda.synchronize()
    t_screen = time.perf_counter() - t0
    err = (mean - E[te]).abs()
    # the method without the network: E(T) ~ c * E(MP2), c fitted on the training set
    emp2 = X[:, -1]
    c = float((emp2[tr] * E[tr]).sum() / (emp2[tr] ** 2).sum())
    err_mp2 = (c * emp2[te] - E[te]).abs()
    acc = spread <= 1e-5
    frac = float(acc.double().mean())
    print(f'\nscreen o{o}v{v}: |E(T)| ~ {float(E.abs().mean()):.3e}, label std '
          f'{float(E.std()):.3e}; test MAE {float(err.mean()):.3e} (MP2 scaling '
          f'{float(err_mp2.mean()):.3e}); accepted {frac:.2f} with max err '
          f'{float(err[acc].max()) if acc.any() else float("nan"):.3e}; exact '
          f'{t_exact / n * 1e3:.2f} ms/instance, screen {t_screen / len(te) * 1e3:.3f} ms/instance')
    n_over = int(((err > 1e-5) & acc).sum())
    print(f'accepted above the 1e-5 Eh gate: {n_over} of {int(acc.sum())}')
    assert float(err.mean()) < 0.5 * float(err_mp2.mean())
    assert frac >= 0.5
    # the spread is an estimate, not a bound: a small measured fraction of the
    # accepted answers may exceed the gate, none by more than a factor of two
    assert n_over <= 0.05 * int(acc.sum())
    assert float(err[acc].max()) <= 2e-5
    exact_cost = t_exact / n * len(te)
    screened_cost = t_screen + (1 - frac) * exact_cost
    assert screened_cost < 0.75 * exact_cost
