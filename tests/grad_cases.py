"""Shared gradient cases: the loss L = sum(g * out) for a fixed random g, its FP64
central differences through the C oracle (h = 1e-6) and FP64 torch autograd of the
dense restatement (oracle/torch_ref.py).  Used by the CPU pinning test and the GPU
adjoint tests."""
import numpy as np
import torch

from oracle import oracle as orc
from oracle import torch_ref as tr

FS = 16000.0
NB = [5, 5, 5]
TMAX, TDIFF = 0.06, 0.025
DIRECTIONAL = dict(mic_pattern="card", orV_rcv=np.array([[0.3, 0.9, 0.1], [-0.5, 0.2, 0.8]]),
                   spkr_pattern="hypcard", orV_src=np.array([[0.7, -0.7, 0.2]]))


def case(seed=0):
    rng = np.random.default_rng(seed)
    L = np.array([4.0, 3.5, 2.7])
    beta = rng.uniform(0.6, 0.9, 6)
    src = 0.5 + rng.uniform(size=(1, 3)) * (L - 1.0)
    rcv = 0.5 + rng.uniform(size=(2, 3)) * (L - 1.0)
    g = rng.standard_normal((1, 2, orc.n_samples(TMAX, FS)))
    return L, beta, src, rcv, g


def loss_oracle(L, beta, src, rcv, g, level=1.0, nb=NB, Tmax=TMAX, Tdiff=TDIFF, **kw):
    y = orc.simulate_rir_f64(L, beta, src, rcv, nb, Tmax, FS, Tdiff=Tdiff, level=level, **kw)
    return float((g * y).sum())


def fd_grads(L, beta, src, rcv, g, h=1e-6, level=1.0, **kw):
    out = {}
    for name, arr in (("beta", beta), ("src", src), ("rcv", rcv)):
        grad = np.zeros_like(arr)
        for idx in np.ndindex(arr.shape):
            p, m = arr.copy(), arr.copy()
            p[idx] += h
            m[idx] -= h
            args = dict(beta=beta, src=src, rcv=rcv)
            args[name] = p
            lp = loss_oracle(L, args["beta"], args["src"], args["rcv"], g, level=level, **kw)
            args[name] = m
            lm = loss_oracle(L, args["beta"], args["src"], args["rcv"], g, level=level, **kw)
            grad[idx] = (lp - lm) / (2 * h)
        out[name] = grad
    lp = loss_oracle(L, beta, src, rcv, g, level=level + h, **kw)
    lm = loss_oracle(L, beta, src, rcv, g, level=level - h, **kw)
    out["level"] = (lp - lm) / (2 * h)
    return out


def autograd_grads(L, beta, src, rcv, g, level=1.0, nb=NB, Tmax=TMAX, Tdiff=TDIFF, **kw):
    b = torch.tensor(beta, requires_grad=True)
    s = torch.tensor(src, requires_grad=True)
    r = torch.tensor(rcv, requires_grad=True)
    lv = torch.tensor(float(level), dtype=torch.float64, requires_grad=True)
    y = tr.simulate(L, b, s, r, nb, Tmax, FS, Tdiff=Tdiff, level=lv, **kw)
    (y * torch.as_tensor(g)).sum().backward()
    glevel = 0.0 if lv.grad is None else float(lv.grad)  # no tail: the output does not depend on level
    return {"beta": b.grad.numpy(), "src": s.grad.numpy(), "rcv": r.grad.numpy(), "level": glevel}
