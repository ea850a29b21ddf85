"""Parity at the full sizes of the batched and trajectory workloads, against the FP64 oracle.

  * 4096 uneven rooms (40,709 RIRs, up to 240,768 images and 16 receivers per room) in one
    autotuned, graph-replayed plan: six rooms checked at 1e-5 relative L2 (early part and
    tail separately, same Philox stream), mixed precision at 1e-3, the table mode against the
    oracle's table mode at 1e-5; replays bit-identical.
  * the moving-source trajectory (128 points x 8 mics, full ISM of 0.5 s, 246,906 images per
    RIR, 10 s of noise): the RIRs of 8 points against the oracle at 1e-5, and the whole
    convolution + float64 overlap-add against the oracle's FP64 overlap-add of the same RIRs
    at 1e-5."""
import numpy as np
import pytest
import torch

from gpurir_amd import torch_ops as T
from gpurir_amd import workloads as W
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def batch_rooms_to_check(rooms, fs):
    """First, most images, longest early part, one with 16 receivers, last."""
    n_img = [int(np.prod(r["nb_img"])) * len(r["pos_rcv"]) for r in rooms]
    n_ism = [orc.n_samples(r["Tdiff"], fs) for r in rooms]
    sixteen = [k for k, r in enumerate(rooms) if len(r["pos_rcv"]) == 16]
    ks = [0, int(np.argmax(n_img)), int(np.argmax(n_ism)), sixteen[len(sixteen) // 2], len(rooms) - 1]
    return list(dict.fromkeys(ks))


@pytest.fixture(scope="module")
def uneven_batch():
    return W.uneven_rooms(n_rooms=4096, seed=0)


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("fp16", 1e-3), ("lut", 1e-5), ("bf16", 1e-2)])
def test_batch_4096_rooms_full_size(uneven_batch, precision, tol):
    rooms, fs = uneven_batch, W.FS
    plan = T.BatchPlan(rooms, fs, precision=precision, seed=1234, call_offset=0)
    assert plan.n_rirs == sum(len(r["pos_rcv"]) for r in rooms)
    tune = plan.autotune()
    print("autotuned:", tune)
    plan.run()
    torch.cuda.synchronize()
    first = plan.out.clone()
    plan.out.fill_(float("nan"))
    plan.run()
    torch.cuda.synchronize()
    assert torch.equal(first, plan.out), "two replays differ"
    for k in batch_rooms_to_check(rooms, fs):
        r = rooms[k]
        o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                             Tdiff=r["Tdiff"], seed=1234, call_offset=0, room_index=k,
                             mode=orc.LUT if precision == "lut" else orc.EXACT_SCATTER)
        g = plan.rirs[k].cpu().numpy()
        n_ism = orc.n_samples(r["Tdiff"], fs)
        e_ism = rel_l2(g[..., :n_ism], o[..., :n_ism])
        e_tail = rel_l2(g[..., n_ism:], o[..., n_ism:])
        print(f"{precision} room {k}: nb {r['nb_img']}, {len(r['pos_rcv'])} receivers, nISM {n_ism}: "
              f"early {e_ism:.2e}, tail {e_tail:.2e}", flush=True)
        assert e_ism <= tol and e_tail <= tol, (k, e_ism, e_tail)


def test_trajectory_full_size():
    cfg = W.trajectory()
    fs = cfg["fs"]
    rirs = T.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"], cfg["Tmax"], fs)
    y = T.trajectory(cfg["signal"], rirs)
    torch.cuda.synchronize()
    assert tuple(rirs.shape) == (128, 8, 8000) and tuple(y.shape) == (160000 + 7999, 8)
    g_rirs = rirs.cpu().numpy()
    pts = [0, 17, 42, 63, 64, 90, 111, 127]
    o = orc.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"][pts], cfg["pos_rcv"], cfg["nb_img"],
                         cfg["Tmax"], fs)
    e_rir = rel_l2(g_rirs[pts], o)
    print(f"trajectory RIRs of {len(pts)} points: rel-L2 {e_rir:.2e}")
    assert e_rir <= 1e-5
    want = orc.simulate_trajectory(cfg["signal"], g_rirs)
    e_ola = rel_l2(y.cpu().numpy(), want)
    print(f"trajectory convolution + overlap-add: rel-L2 {e_ola:.2e}")
    assert e_ola <= 1e-5
