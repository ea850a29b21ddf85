"""Pins the gradient definition on the CPU: FP64 central differences of the C oracle
(h = 1e-6) against FP64 torch autograd of an independent dense restatement.
The GPU adjoint is then checked against both (test_gpu_adjoint.py)."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from oracle import torch_ref as tr

FS = 16000.0


def case(seed=0):
    rng = np.random.default_rng(seed)
    L = np.array([4.0, 3.5, 2.7])
    beta = rng.uniform(0.6, 0.9, 6)
    src = 0.5 + rng.uniform(size=(1, 3)) * (L - 1.0)
    rcv = 0.5 + rng.uniform(size=(2, 3)) * (L - 1.0)
    g = rng.standard_normal((1, 2, orc.n_samples(0.06, FS)))
    return L, beta, src, rcv, g


def loss_oracle(L, beta, src, rcv, g, level=1.0, **kw):
    y = orc.simulate_rir_f64(L, beta, src, rcv, [5, 5, 5], 0.06, FS, Tdiff=0.025, level=level, **kw)
    return float((g * y).sum())


def fd_grads(L, beta, src, rcv, g, h=1e-6, **kw):
    out = {}
    for name, arr in (("beta", beta), ("src", src), ("rcv", rcv)):
        grad = np.zeros_like(arr)
        for idx in np.ndindex(arr.shape):
            p, m = arr.copy(), arr.copy()
            p[idx] += h
            m[idx] -= h
            args = dict(beta=beta, src=src, rcv=rcv)
            args[name] = p
            lp = loss_oracle(L, args["beta"], args["src"], args["rcv"], g, **kw)
            args[name] = m
            lm = loss_oracle(L, args["beta"], args["src"], args["rcv"], g, **kw)
            grad[idx] = (lp - lm) / (2 * h)
        out[name] = grad
    lp = loss_oracle(L, beta, src, rcv, g, level=1.0 + h, **kw)
    lm = loss_oracle(L, beta, src, rcv, g, level=1.0 - h, **kw)
    out["level"] = (lp - lm) / (2 * h)
    return out


def autograd_grads(L, beta, src, rcv, g, **kw):
    b = torch.tensor(beta, requires_grad=True)
    s = torch.tensor(src, requires_grad=True)
    r = torch.tensor(rcv, requires_grad=True)
    lv = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    y = tr.simulate(L, b, s, r, [5, 5, 5], 0.06, FS, Tdiff=0.025, level=lv, **kw)
    (y * torch.as_tensor(g)).sum().backward()
    return {"beta": b.grad.numpy(), "src": s.grad.numpy(), "rcv": r.grad.numpy(), "level": float(lv.grad)}


@pytest.mark.parametrize("patterns", [dict(), dict(mic_pattern="card", orV_rcv=np.array([[0.3, 0.9, 0.1],
                                                                                           [-0.5, 0.2, 0.8]]),
                                                  spkr_pattern="hypcard", orV_src=np.array([[0.7, -0.7, 0.2]]))])
def test_finite_differences_match_autograd(patterns):
    L, beta, src, rcv, g = case()
    fd = fd_grads(L, beta, src, rcv, g, **patterns)
    ag = autograd_grads(L, beta, src, rcv, g, **patterns)
    for k in ("beta", "src", "rcv"):
        np.testing.assert_allclose(ag[k], fd[k], rtol=1e-6, atol=1e-7 * np.abs(fd[k]).max())
    assert abs(ag["level"] - fd["level"]) <= 1e-6 * abs(fd["level"])
