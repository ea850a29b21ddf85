"""Gradients of the 256-room inverse problem at its own size, against FP64 central
differences (h = 1e-6) of the C oracle.

Workload: 256 rooms (4-8 x 3-6 x 2.5-3.5 m, T60 0.3-0.8 s, 1 source x 4 omni receivers,
fs 16 kHz, Tmax 0.5 s, Tdiff at 20 dB); the loss is the MSE between the simulated RIRs
and targets simulated from betas scaled by 1 + U(-0.2, 0.2) and sources moved by
N(0, 0.1^2) m.  The adjoint runs with the default (basis) method over every 768-delay
tile of each RIR (up to 6 tiles, 122,472 images per room).

Bar: every gradient group (6 betas, source xyz, receiver xyz, level) of every checked
room within 1e-4 relative L2 of the FP64 finite differences.  The target is the GPU's
own simulation of the perturbed rooms, held fixed; the FP64 loss uses it as given."""
import numpy as np
import pytest
import torch

from gpurir_amd import torch_ops as T
from gpurir_amd import workloads as W
from oracle import grad_check as gc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-4


def checked_rooms(rooms, fs):
    """The first and last rooms, and those with the most images, the fewest images and the
    longest early part."""
    n_img = [int(np.prod(r["nb_img"])) for r in rooms]
    n_ism = [orc.n_samples(r["Tdiff"], fs) for r in rooms]
    ks = [0, int(np.argmax(n_img)), int(np.argmax(n_ism)), int(np.argmin(n_img)), len(rooms) - 1]
    return list(dict.fromkeys(ks))


def gpu_room_grads(gb, gs, gr, gl, rooms, k):
    so = sum(len(r["pos_src"]) for r in rooms[:k])
    ro = sum(len(r["pos_rcv"]) for r in rooms[:k])
    ns, nr = len(rooms[k]["pos_src"]), len(rooms[k]["pos_rcv"])
    return {"beta": gb[k].double().cpu().numpy(), "src": gs[so:so + ns].double().cpu().numpy(),
            "rcv": gr[ro:ro + nr].double().cpu().numpy(), "level": float(gl[k])}


def test_gradient_workload_matches_finite_differences():
    fs = W.FS
    rooms = W.gradient_rooms(256, seed=0)
    target = T.BatchPlan(W.gradient_targets(rooms, seed=1), fs, seed=1234).run().clone()
    plan = T.BatchPlan(rooms, fs, seed=1234)
    y = plan.run()
    scale = 1.0 / y.numel()
    gb, gs, gr, gl = plan.adjoint(2.0 * scale * (y - target))
    torch.cuda.synchronize()
    t_host = target.cpu().numpy()
    off = 0
    offsets = []
    for r in rooms:
        offsets.append(off)
        off += len(r["pos_src"]) * len(r["pos_rcv"]) * orc.n_samples(r["Tmax"], fs)
    worst = 0.0
    for k in checked_rooms(rooms, fs):
        r = rooms[k]
        shape = (len(r["pos_src"]), len(r["pos_rcv"]), orc.n_samples(r["Tmax"], fs))
        tk = t_host[offsets[k]:offsets[k] + int(np.prod(shape))].reshape(shape)
        want = gc.room_mse_grads(r, tk, scale, fs, seed=1234, room_index=k)
        got = gpu_room_grads(gb, gs, gr, gl, rooms, k)
        errs = gc.group_errors(got, want)
        print(f"room {k} (nb {r['nb_img']}, nISM {orc.n_samples(r['Tdiff'], fs)}): "
              + ", ".join(f"d/d{g} {e:.2e}" for g, e in errs.items()), flush=True)
        for g, e in errs.items():
            assert e <= TOL, f"room {k} d/d{g}: rel-L2 {e:.3e} (gpu {got[g]}, fd {want[g]})"
        worst = max(worst, max(errs.values()))
    print(f"largest rel-L2 against finite differences: {worst:.3e}")


@pytest.mark.parametrize("method,wide", [("basis", False), ("taps", False), ("basis", True)],
                         ids=["basis", "taps", "basis-long-axis-tables"])
def test_multi_tile_adjoint_small_room(method, wide):
    """1600 early samples per RIR (Tdiff 0.1 s at 16 kHz, every image of the early part): the
    basis adjoint's delay tiles (two of 768 delays and a last one that runs to the last delay
    with a tap in the early part) and the tap adjoint's tiles, against FP64 central differences
    of the C oracle.  With 100 images per axis (wide: 300 axis-table entries, of which the early
    part uses about a third) the basis adjoint's LDS no longer fits two workgroups per CU with
    768-delay tiles and it takes tiles of 512."""
    fs = 16000.0
    L = np.array([4.0, 3.5, 2.7], np.float32)
    rng = np.random.default_rng(11)
    nb = [100, 100, 100] if wide else [int(np.ceil(2 * 0.1 / (x / 343.0))) for x in L]
    room = dict(room_sz=L, nb_img=nb, Tmax=0.13, Tdiff=0.1,
                beta=rng.uniform(0.7, 0.95, 6).astype(np.float32),
                pos_src=(0.5 + rng.uniform(size=(1, 3)) * (L - 1.0)).astype(np.float32),
                pos_rcv=(0.5 + rng.uniform(size=(2, 3)) * (L - 1.0)).astype(np.float32), level=0.8)
    assert orc.n_samples(room["Tdiff"], fs) == 1600  # 2 x 768 + 64
    dev = torch.device("cuda", 0)
    b = torch.tensor(room["beta"], device=dev, requires_grad=True)
    s = torch.tensor(room["pos_src"], device=dev, requires_grad=True)
    rr = torch.tensor(room["pos_rcv"], device=dev, requires_grad=True)
    lv = torch.tensor(0.8, dtype=torch.float32, device=dev, requires_grad=True)
    y = T.simulate_rir_diff(L, b, s, rr, room["nb_img"], room["Tmax"], fs, Tdiff=room["Tdiff"], level=lv,
                            method=method)
    tgt = np.random.default_rng(12).standard_normal(tuple(y.shape)) * 0.01
    ((y - torch.as_tensor(tgt, dtype=torch.float32, device=dev)) ** 2).sum().backward()
    torch.cuda.synchronize()
    got = {"beta": b.grad.double().cpu().numpy(), "src": s.grad.double().cpu().numpy(),
           "rcv": rr.grad.double().cpu().numpy(), "level": float(lv.grad)}
    want = gc.room_mse_grads(dict(room, level=np.float32(0.8)), tgt, 1.0, fs)
    errs = gc.group_errors(got, want)
    print(method, "wide" if wide else "", errs)
    for g, e in errs.items():
        assert e <= TOL, f"{method} d/d{g}: rel-L2 {e:.3e}"


def test_diff_plan_optimiser_loop_tracks_changing_parameters():
    """An SGD loop through DiffPlan (parameters in device buffers read by the captured graph):
    after every update the output and the gradients equal those of a plan built from scratch
    with the updated parameters, bit for bit."""
    fs = W.FS
    rooms = W.gradient_rooms(6, seed=4)
    dev = torch.device("cuda", 0)
    target = T.BatchPlan(W.gradient_targets(rooms, seed=5), fs, seed=1234).run().clone()
    dp = T.DiffPlan(rooms, fs, seed=1234)
    beta = torch.tensor(np.stack([r["beta"] for r in rooms]), device=dev, requires_grad=True)
    src = torch.tensor(np.concatenate([r["pos_src"] for r in rooms]), device=dev, requires_grad=True)
    rcv = torch.tensor(np.concatenate([r["pos_rcv"] for r in rooms]), device=dev, requires_grad=True)
    level = torch.ones(len(rooms), device=dev, requires_grad=True)
    opt = torch.optim.SGD([beta, src, level], lr=1e3)
    for step in range(3):
        opt.zero_grad()
        rcv.grad = None  # receivers are differentiated but not optimised
        y = dp(beta, src, rcv, level)
        g = (2.0 / y.numel()) * (y.detach() - target)  # d MSE / dy, computed as for the fresh plan below
        y.backward(g)
        fresh_rooms = [dict(r, beta=beta[k].detach().cpu().numpy(), pos_src=src[k:k + 1].detach().cpu().numpy(),
                            level=float(level[k].detach())) for k, r in enumerate(rooms)]
        fresh = T.BatchPlan(fresh_rooms, fs, seed=1234)
        yf = fresh.run()
        assert torch.equal(yf, y), f"step {step}: output"
        gb, gs, gr, gl = fresh.adjoint((2.0 / yf.numel()) * (yf - target))
        assert torch.equal(gb, beta.grad) and torch.equal(gs, src.grad) and torch.equal(gr, rcv.grad), step
        assert torch.equal(gl, level.grad), step
        opt.step()
        with torch.no_grad():
            beta.clamp_(0.0, 1.0)
        assert float(beta.grad.abs().max()) > 0.0


def test_keep_outputs_backward_out_of_order():
    """keep=True: two forward passes, then the first one's backward, gives the first pass's
    gradients (its output and parameters are restored)."""
    fs = W.FS
    rooms = W.gradient_rooms(3, seed=6)
    dev = torch.device("cuda", 0)
    dp = T.DiffPlan(rooms, fs)
    b0 = torch.tensor(np.stack([r["beta"] for r in rooms]), device=dev)
    s0 = torch.tensor(np.concatenate([r["pos_src"] for r in rooms]), device=dev)
    r0 = torch.tensor(np.concatenate([r["pos_rcv"] for r in rooms]), device=dev)
    l0 = torch.ones(3, device=dev)
    g = torch.randn(dp.out.shape, generator=torch.Generator().manual_seed(3)).to(dev)
    want = []
    for scale in (1.0, 0.9):
        b = (b0 * scale).requires_grad_(True)
        y = dp(b, s0, r0, l0, keep=True)
        (y * g).sum().backward()
        want.append(b.grad.clone())
    b1 = (b0 * 1.0).requires_grad_(True)
    b2 = (b0 * 0.9).requires_grad_(True)
    y1 = dp(b1, s0, r0, l0, keep=True)
    y2 = dp(b2, s0, r0, l0, keep=True)
    (y1 * g).sum().backward()
    (y2 * g).sum().backward()
    assert torch.equal(b1.grad, want[0]) and torch.equal(b2.grad, want[1])
    b3 = b0.clone().requires_grad_(True)
    y3 = dp(b3, s0, r0, l0)
    dp(b0, s0, r0, l0)
    with pytest.raises(RuntimeError):
        (y3 * g).sum().backward()
