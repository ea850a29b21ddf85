"""The hand-written adjoint kernels (backward pass) against FP64 central differences
of the C oracle (h = 1e-6) and FP64 torch autograd of the dense restatement.

Bar: for every gradient group (beta, source positions, receiver positions, level) the
relative L2 error is <= 1e-4.  Inputs are rounded to float32 first so that the
FP32 GPU path and the FP64 references see the same parameters."""
import numpy as np
import pytest
import torch

from gpurir_amd import torch_ops as T
from grad_cases import DIRECTIONAL, FS, NB, TDIFF, TMAX, autograd_grads, case, fd_grads
from oracle import oracle as orc
from oracle import torch_ref as tr

pytestmark = pytest.mark.gpu
TOL = 1e-4
GROUPS = ("beta", "src", "rcv", "level")


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_grads(L, beta, src, rcv, g, level=1.0, nb=NB, Tmax=TMAX, Tdiff=TDIFF, fs=FS, chunk_images=0,
              tile_samples=0, **kw):
    dev = torch.device("cuda", 0)
    b = torch.tensor(beta, dtype=torch.float32, device=dev, requires_grad=True)
    s = torch.tensor(src, dtype=torch.float32, device=dev, requires_grad=True)
    r = torch.tensor(rcv, dtype=torch.float32, device=dev, requires_grad=True)
    lv = torch.tensor(float(level), dtype=torch.float32, device=dev, requires_grad=True)
    y = T.simulate_rir_diff(L, b, s, r, nb, Tmax, fs, Tdiff=Tdiff, level=lv, chunk_images=chunk_images,
                            tile_samples=tile_samples, **kw)
    (y * torch.as_tensor(g, dtype=torch.float32, device=dev)).sum().backward()
    torch.cuda.synchronize()
    return {"beta": b.grad.double().cpu().numpy(), "src": s.grad.double().cpu().numpy(),
            "rcv": r.grad.double().cpu().numpy(), "level": float(lv.grad)}


def assert_close(got, want, tol=TOL, what=""):
    for k in GROUPS:
        e = rel(got[k], want[k]) if np.any(np.asarray(want[k]) != 0) else float(np.abs(got[k]).max())
        print(f"{what} d/d{k}: rel-L2 {e:.3e}")
        assert e <= tol, f"{what} d/d{k}: rel-L2 {e:.3e} (gpu {got[k]}, ref {want[k]})"


def rounded_case(seed=0):
    L, beta, src, rcv, g = case(seed)
    return f32(L), f32(beta), f32(src), f32(rcv), g


@pytest.mark.parametrize("method", ["taps", "basis"])
@pytest.mark.parametrize("patterns", [dict(), DIRECTIONAL], ids=["omni", "card-hypcard"])
def test_adjoint_matches_fd_and_autograd(patterns, method):
    """Both accumulation methods: the tap scatter and the 10-term basis adjoint (the forward of
    the basis method is within 1e-6 of the exact sinc, its adjoint's h' within 5e-7 per tap)."""
    L, beta, src, rcv, g = rounded_case()
    kw = {k: (f32(v) if k.startswith("orV") else v) for k, v in patterns.items()}
    got = gpu_grads(L, beta, src, rcv, g, method=method, **kw)
    assert_close(got, fd_grads(L, beta, src, rcv, g, **kw), what="vs finite differences")
    assert_close(got, autograd_grads(L, beta, src, rcv, g, **kw), what="vs autograd")


def test_adjoint_split_items_agree():
    """Many image chunks and time tiles per RIR give the same gradients."""
    L, beta, src, rcv, g = rounded_case(1)
    a = gpu_grads(L, beta, src, rcv, g)
    b = gpu_grads(L, beta, src, rcv, g, chunk_images=37, tile_samples=100)
    assert_close(b, a, tol=1e-6, what="split vs whole")
    assert_close(b, autograd_grads(L, beta, src, rcv, g), what="split vs autograd")


def test_adjoint_zero_wall_and_level():
    """beta = 0 on one wall (att = 0 for its images, derivative not 0) and level != 1."""
    L, beta, src, rcv, g = rounded_case(2)
    beta[3] = 0.0
    got = gpu_grads(L, beta, src, rcv, g, level=0.7)
    assert_close(got, fd_grads(L, beta, src, rcv, g, level=0.7), what="vs finite differences")


def test_adjoint_without_tail():
    L, beta, src, rcv, g = rounded_case(3)
    g = g[..., :orc.n_samples(TDIFF, FS)]
    got = gpu_grads(L, beta, src, rcv, g, Tmax=TDIFF, Tdiff=TDIFF)
    want = autograd_grads(L, beta, src, rcv, g, Tmax=TDIFF, Tdiff=TDIFF)
    assert got["level"] == 0.0
    for k in ("beta", "src", "rcv"):
        assert rel(got[k], want[k]) <= TOL


def test_adjoint_192k():
    """fs = 192 kHz: a window of 1536 taps (the tap adjoint; the automatic choice)."""
    L, beta, src, rcv, _ = rounded_case(7)
    fs = 192000.0
    nS = orc.n_samples(0.012, fs)
    g = np.random.default_rng(8).standard_normal((1, 2, nS))
    kw = dict(nb=[3, 3, 3], Tmax=0.012, Tdiff=0.008)
    got = gpu_grads(L, beta, src, rcv, g, fs=fs, **kw)
    b = torch.tensor(beta, requires_grad=True)
    s = torch.tensor(src, requires_grad=True)
    r = torch.tensor(rcv, requires_grad=True)
    lv = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    y = tr.simulate(L, b, s, r, kw["nb"], kw["Tmax"], fs, Tdiff=kw["Tdiff"], level=lv)
    (y * torch.as_tensor(g)).sum().backward()
    want = {"beta": b.grad.numpy(), "src": s.grad.numpy(), "rcv": r.grad.numpy(), "level": float(lv.grad)}
    assert_close(got, want, what="192 kHz vs autograd")


@pytest.mark.parametrize("method", ["taps", "basis"])
def test_adjoint_odd_window_44k1(method):
    """fs = 44.1 kHz: Tw = 353 (odd): the masked tap adjoint, and the basis adjoint."""
    L, beta, src, rcv, _ = rounded_case(4)
    fs = 44100.0
    nS = orc.n_samples(0.02, fs)
    g = np.random.default_rng(5).standard_normal((1, 2, nS))
    kw = dict(nb=[3, 3, 3], Tmax=0.02, Tdiff=0.012)
    got = gpu_grads(L, beta, src, rcv, g, fs=fs, method=method, **kw)
    b = torch.tensor(beta, requires_grad=True)
    s = torch.tensor(src, requires_grad=True)
    r = torch.tensor(rcv, requires_grad=True)
    lv = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    y = tr.simulate(L, b, s, r, kw["nb"], kw["Tmax"], fs, Tdiff=kw["Tdiff"], level=lv)
    (y * torch.as_tensor(g)).sum().backward()
    want = {"beta": b.grad.numpy(), "src": s.grad.numpy(), "rcv": r.grad.numpy(), "level": float(lv.grad)}
    assert_close(got, want, what="44.1 kHz vs autograd")


@pytest.mark.parametrize("method", ["taps", "basis"])
def test_adjoint_is_deterministic(method):
    L, beta, src, rcv, g = rounded_case(6)
    a = gpu_grads(L, beta, src, rcv, g, chunk_images=64, method=method)
    b = gpu_grads(L, beta, src, rcv, g, chunk_images=64, method=method)
    for k in GROUPS:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def batch_case(n=4, seed=7):
    rng = np.random.default_rng(seed)
    rooms, betas, ps, pr, lv = [], [], [], [], []
    for k in range(n):
        L = f32(rng.uniform([3.0, 3.0, 2.5], [5.0, 4.5, 3.2]))
        nr = int(rng.integers(1, 4))
        rooms.append(dict(room_sz=L, nb_img=[int(x) for x in rng.integers(3, 6, 3)], Tmax=float(rng.uniform(0.03, 0.06)),
                          Tdiff=float(rng.uniform(0.015, 0.025)), n_src=1, n_rcv=nr))
        betas.append(f32(rng.uniform(0.5, 0.9, 6)))
        ps.append(f32(0.5 + rng.uniform(size=(1, 3)) * (L - 1.0)))
        pr.append(f32(0.5 + rng.uniform(size=(nr, 3)) * (L - 1.0)))
        lv.append(float(np.float32(rng.uniform(0.5, 1.5))))
    return rooms, betas, ps, pr, lv


@pytest.mark.parametrize("method", ["taps", "basis"])
def test_batch_adjoint_uneven_rooms(method):
    rooms, betas, ps, pr, lv = batch_case()
    dev = torch.device("cuda", 0)
    b = torch.tensor(np.stack(betas), dtype=torch.float32, device=dev, requires_grad=True)
    s = torch.tensor(np.concatenate(ps), dtype=torch.float32, device=dev, requires_grad=True)
    r = torch.tensor(np.concatenate(pr), dtype=torch.float32, device=dev, requires_grad=True)
    l = torch.tensor(lv, dtype=torch.float32, device=dev, requires_grad=True)
    out, views = T.simulate_batch_diff(rooms, b, s, r, l, FS, room_index=3, method=method)
    rng = np.random.default_rng(8)
    gs = [rng.standard_normal(tuple(v.shape)) for v in views]
    loss = sum((v * torch.as_tensor(gk, dtype=torch.float32, device=dev)).sum() for v, gk in zip(views, gs))
    loss.backward()
    torch.cuda.synchronize()
    so = ro = 0
    for k, room in enumerate(rooms):
        nr = room["n_rcv"]
        bb = torch.tensor(betas[k], requires_grad=True)
        ss = torch.tensor(ps[k], requires_grad=True)
        rr = torch.tensor(pr[k], requires_grad=True)
        ll = torch.tensor(lv[k], dtype=torch.float64, requires_grad=True)
        y = tr.simulate(room["room_sz"], bb, ss, rr, room["nb_img"], room["Tmax"], FS, Tdiff=room["Tdiff"],
                        level=ll, room_index=3 + k)
        (y * torch.as_tensor(gs[k])).sum().backward()
        got = {"beta": b.grad[k].double().cpu().numpy(), "src": s.grad[so:so + 1].double().cpu().numpy(),
               "rcv": r.grad[ro:ro + nr].double().cpu().numpy(), "level": float(l.grad[k])}
        want = {"beta": bb.grad.numpy(), "src": ss.grad.numpy(), "rcv": rr.grad.numpy(), "level": float(ll.grad)}
        assert_close(got, want, what=f"room {k}")
        so += 1
        ro += nr


@pytest.mark.parametrize("method", ["taps", "basis"])
def test_batch_plan_adjoint_matches_autograd_path(method):
    rooms, betas, ps, pr, lv = batch_case(seed=9)
    dev = torch.device("cuda", 0)
    plan_rooms = [dict(r, beta=betas[k], pos_src=ps[k], pos_rcv=pr[k], level=lv[k]) for k, r in enumerate(rooms)]
    plan = T.BatchPlan(plan_rooms, FS, method=method)
    y = plan.run()
    g = torch.randn(y.shape, generator=torch.Generator().manual_seed(1)).to(dev)
    gb, gs, gr, gl = plan.adjoint(g)
    b = torch.tensor(np.stack(betas), dtype=torch.float32, device=dev, requires_grad=True)
    s = torch.tensor(np.concatenate(ps), dtype=torch.float32, device=dev, requires_grad=True)
    r = torch.tensor(np.concatenate(pr), dtype=torch.float32, device=dev, requires_grad=True)
    l = torch.tensor(lv, dtype=torch.float32, device=dev, requires_grad=True)
    out, _ = T.simulate_batch_diff(rooms, b, s, r, l, FS, method=method)
    assert torch.equal(out, y)
    (out * g).sum().backward()
    torch.cuda.synchronize()
    for got, want in ((gb, b.grad), (gs, s.grad), (gr, r.grad), (gl, l.grad)):
        assert torch.equal(got, want)
