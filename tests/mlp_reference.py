"""CPU reference of the BF16 MLP (test infrastructure).

The same network as pyscfad_amd/screen/mlp.py, written with torch float32 on
the CPU and differentiated by torch autograd: weights are the bf16 copies the
GPU uses, inputs are the bf16 inputs, hidden activations are rounded to bf16
in the forward pass with a straight-through gradient (the GPU backward uses
SiLU'(z) at the FP32 pre-activation), and the loss is the mean squared error.
"""
import torch


def bf16_round(x):
    return x.to(torch.bfloat16).to(torch.float32)


def bits_to_f32(b):
    """bf16 bit patterns held in an int16 tensor -> float32 values."""
    return (b.to(torch.int32) & 0xFFFF).bitwise_left_shift(16).view(torch.float32)


def forward(Ws, bs, x):
    """Ws[l] (out, in), bs[l] (out,), x (batch, in); all float32 CPU tensors."""
    h = x
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ W.T + b
        if l < len(Ws) - 1:
            y = torch.nn.functional.silu(z)
            h = y + (bf16_round(y) - y).detach()
        else:
            h = z
    return h


def loss_and_grads(Ws, bs, x, y):
    Ws = [W.clone().requires_grad_(True) for W in Ws]
    bs = [b.clone().requires_grad_(True) for b in bs]
    pred = forward(Ws, bs, x)
    loss = ((pred - y) ** 2).mean()
    loss.backward()
    return pred.detach(), loss.detach(), [(W.grad, b.grad) for W, b in zip(Ws, bs)]


def adam(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step with decoupled weight decay, in float64 for reference."""
    w, g, m1, m2 = (t.double() for t in (w, g, m1, m2))
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    w = w - lr * ((m1 / c1) / ((m2 / c2).sqrt() + eps) + wd * w)
    return w, m1, m2
