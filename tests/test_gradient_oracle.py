"""Pins the gradient definition on the CPU: FP64 central differences of the C oracle
(h = 1e-6) against FP64 torch autograd of an independent dense restatement.
The GPU adjoint is then checked against both (test_gpu_adjoint.py)."""
import numpy as np
import pytest

from grad_cases import DIRECTIONAL, autograd_grads, case, fd_grads


@pytest.mark.parametrize("patterns", [dict(), DIRECTIONAL])
def test_finite_differences_match_autograd(patterns):
    L, beta, src, rcv, g = case()
    fd = fd_grads(L, beta, src, rcv, g, **patterns)
    ag = autograd_grads(L, beta, src, rcv, g, **patterns)
    for k in ("beta", "src", "rcv"):
        np.testing.assert_allclose(ag[k], fd[k], rtol=1e-6, atol=1e-7 * np.abs(fd[k]).max())
    assert abs(ag["level"] - fd["level"]) <= 1e-6 * abs(fd["level"])
