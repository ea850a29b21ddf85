"""The no-allocation entry point (gpurir_simulate_rir_device_noalloc) against the entry point
it is a variant of (gpurir_simulate_rir_device), through the C ABI: the same plan and kernels,
so the same bits, on calls in a row whose inputs shrink (its scratch array keeps the previous
call's contents), by both accumulation methods; and a launch too large for its scratch is
rejected with a message."""
import ctypes

import numpy as np
import pytest
import torch

from gpurir_amd import _lib
from gpurir_amd import torch_ops as T
from gpurir_amd import workloads as W

pytestmark = pytest.mark.gpu


def call(name, cfg, method, out, seed=1234, call_offset=0):
    dev = out.device
    ps = torch.as_tensor(np.asarray(cfg["pos_src"], np.float32), device=dev)
    pr = torch.as_tensor(np.asarray(cfg["pos_rcv"], np.float32), device=dev)
    os_ = torch.zeros_like(ps) if cfg["orV_src"] is None else torch.as_tensor(cfg["orV_src"], device=dev)
    or_ = torch.zeros_like(pr) if cfg["orV_rcv"] is None else torch.as_tensor(cfg["orV_rcv"], device=dev)
    room = np.ascontiguousarray(np.asarray(cfg["room_sz"], np.float32))
    beta = np.ascontiguousarray(np.asarray(cfg["beta"], np.float32))
    nb = np.ascontiguousarray(np.asarray(cfg["nb_img"], np.int32))
    o = T.options(seed=seed, call_offset=call_offset, method=method)
    pat = {"omni": 0, "card": 2, "hypcard": 3}
    rc = getattr(_lib.load(), name)(_lib.fptr(room), _lib.fptr(beta), ps.data_ptr(), ps.shape[0], pr.data_ptr(),
                                    pr.shape[0], os_.data_ptr(), or_.data_ptr(), pat[cfg["spkr_pattern"]],
                                    pat[cfg["mic_pattern"]], _lib.iptr(nb), float(cfg["Tdiff"]), float(cfg["Tmax"]),
                                    float(cfg["fs"]), 343.0, ctypes.byref(o), out.data_ptr(),
                                    torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return out


def shape(cfg):
    return (len(cfg["pos_src"]), len(cfg["pos_rcv"]), T.n_samples(cfg["Tmax"], cfg["fs"]))


@pytest.mark.parametrize("method", ["basis", "taps"])
def test_noalloc_matches_the_allocating_entry_point(method):
    dev = torch.device("cuda", 0)
    # the mid room (128 RIRs; by the tap method ~2048 items, more than the resident workgroups,
    # taken off the work counters in the scratch array, and 26 MB of partial rows), then the
    # smaller plain room (2 RIRs of 82,369 images: 161 image chunks each by the tap method, one
    # workgroup per item)
    for k, cfg in enumerate([W.mid_room(), W.plain_room()]):
        want = call("gpurir_simulate_rir_device", cfg, method, torch.empty(shape(cfg), device=dev), call_offset=k)
        got = torch.full(shape(cfg), float("nan"), device=dev)
        call("gpurir_simulate_rir_device_noalloc", cfg, method, got, call_offset=k)
        assert torch.equal(got, want), f"call {k} ({method})"
        again = torch.full(shape(cfg), float("nan"), device=dev)
        call("gpurir_simulate_rir_device_noalloc", cfg, method, again, call_offset=k)
        assert torch.equal(again, want), f"repeated call {k} ({method})"


def test_noalloc_rejects_launches_beyond_its_scratch():
    dev = torch.device("cuda", 0)
    cfg = W.trajectory(n_pts=128)
    cfg.update(spkr_pattern="omni", mic_pattern="omni", orV_src=None, orV_rcv=None)
    out = torch.empty(shape(cfg), device=dev)
    with pytest.raises(ValueError, match="no-allocation entry point has"):
        call("gpurir_simulate_rir_device_noalloc", cfg, "taps", out)
