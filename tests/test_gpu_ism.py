"""GPU parity of the image-source kernels (stages a, b, c) against the FP64 oracle.

Tolerances (relative L2 against FP64): FP32 exact sinc 1e-5 on the early part and
on the tail (same Philox noise); the LUT kernel against the oracle's LUT mode 1e-5.
Integer image offsets n0 must be identical."""
import numpy as np
import pytest
import torch

import gpurir_amd as gpuRIR
from gpurir_amd import torch_ops as T
from gpurir_amd import workloads as W
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_both(cfg, precision="fp32", seed=1234, call_offset=0, **kw):
    args = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
                nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"],
                spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"], orV_src=cfg["orV_src"],
                orV_rcv=cfg["orV_rcv"])
    g = T.simulate_rir(**args, precision=precision, seed=seed, call_offset=call_offset, **kw)
    torch.cuda.synchronize()
    mode = orc.LUT if precision == "lut" else orc.EXACT_SCATTER
    o = orc.simulate_rir(**args, mode=mode, seed=seed, call_offset=call_offset)
    return g.cpu().numpy(), o


def check_parity(g, o, nISM, tol=1e-5):
    e_ism = rel_l2(g[..., :nISM], o[..., :nISM])
    assert e_ism <= tol, f"ISM rel-L2 {e_ism:.3e}"
    if o.shape[-1] > nISM:
        e_tail = rel_l2(g[..., nISM:], o[..., nISM:])
        assert e_tail <= tol, f"tail rel-L2 {e_tail:.3e}"
    return e_ism


def test_plain_room_fp32():
    cfg = W.plain_room()
    g, o = run_both(cfg)
    assert g.shape == (1, 2, 12800)
    check_parity(g, o, 2800)


def test_mid_room_fp32_directional():
    cfg = W.mid_room()
    g, o = run_both(cfg)
    assert g.shape == (4, 32, 17600)
    check_parity(g, o, 3200)


def test_mid_room_lut_against_oracle_lut():
    cfg = W.mid_room(n_src=2, n_rcv=8)
    g, o = run_both(cfg, precision="lut")
    check_parity(g, o, 3200)
    # and the table's own error against the exact sinc is stated, not gated
    ex = orc.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"], cfg["Tmax"],
                          cfg["fs"], Tdiff=cfg["Tdiff"], spkr_pattern=cfg["spkr_pattern"],
                          mic_pattern=cfg["mic_pattern"], orV_src=cfg["orV_src"], orV_rcv=cfg["orV_rcv"])
    assert rel_l2(g[..., :3200], ex[..., :3200]) < 5e-3


def test_mid_room_lut_full_size():
    """The reference's default mode on the whole mid room (4 x 32 directional RIRs, 896 tiles:
    the automatic choice takes the 16 phase rows on the matrix cores) against the oracle's
    table mode, and the tap scatter of the same table."""
    cfg = W.mid_room()
    g, o = run_both(cfg, precision="lut")
    g_t, _ = run_both(cfg, precision="lut", method="taps")
    e = check_parity(g, o, 3200)
    e_t = check_parity(g_t, o, 3200)
    print(f"mid room LUT: phase rows {e:.2e}, tap scatter {e_t:.2e}")


@pytest.mark.parametrize("cfg_name", ["plain", "mid"])
def test_image_offsets_identical(cfg_name):
    cfg = W.plain_room() if cfg_name == "plain" else W.mid_room(n_src=2, n_rcv=4)
    kw = dict(spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"], orV_src=cfg["orV_src"],
              orV_rcv=cfg["orV_rcv"])
    n0, fr, amp = T.image_list(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"], cfg["fs"],
                               **kw)
    on0, ofr, oamp, _ = orc.images(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"],
                                   cfg["fs"], **kw)
    n0 = n0.cpu().numpy()
    assert np.array_equal(n0.astype(np.int64), on0), f"{np.count_nonzero(n0 != on0)} mismatched n0"
    assert np.max(np.abs(fr.cpu().numpy() - ofr)) < 1e-7
    a = amp.cpu().numpy()
    assert np.max(np.abs(a - oamp)) <= 2e-6 * np.max(np.abs(oamp))


# 8 kHz: Tw = 64 (1 pass of 64 lanes), 16 kHz: 128 (2), 22.05 kHz: 176 (3), 32 kHz: 256 (4),
# 44.1 kHz: 353 (odd, 6 with masked lanes), 48 kHz: 384 (6), 96 kHz: 768 (12), 176.4 kHz: 1411
# (odd, two groups of 12 passes), 192 kHz: 1536 (two groups)
ALL_FS = [8000.0, 16000.0, 22050.0, 32000.0, 44100.0, 48000.0, 96000.0, 176400.0, 192000.0]


@pytest.mark.parametrize("fs", ALL_FS)
def test_sampling_rates_odd_and_even_windows(fs):
    cfg = W.plain_room()
    cfg["fs"] = fs
    cfg["Tmax"] = 0.12
    cfg["Tdiff"] = 0.06
    cfg["nb_img"] = gpuRIR.t2n(0.06, cfg["room_sz"])
    g, o = run_both(cfg)
    check_parity(g, o, orc.n_samples(0.06, fs))
    g1, _ = run_both(cfg, method="taps")
    check_parity(g1, o, orc.n_samples(0.06, fs))
    g2, o2 = run_both(cfg, precision="lut")
    check_parity(g2, o2, orc.n_samples(0.06, fs))
    if fs <= 64000.0:  # the table mode on the matrix-core filters (16 phase rows)
        g3, _ = run_both(cfg, precision="lut", method="basis")
        check_parity(g3, o2, orc.n_samples(0.06, fs))


def test_basis_and_taps_methods():
    """FP32: the Chebyshev-basis accumulation and the tap scatter both meet the 1e-5 bar on a
    directional room with several sources; the basis error is dominated by its
    interpolation (<= 5e-7 per tap of 1), so the two agree to a few 1e-6.  The automatic
    choice takes the basis method for this launch (48 RIRs x 7 tiles of 512 samples) and
    the tap scatter for a launch of fewer than 256 tiles (one room, two receivers)."""
    cfg = W.mid_room(n_src=3, n_rcv=16, seed=5)
    g_b, o = run_both(cfg, method="basis")
    g_t, _ = run_both(cfg, method="taps")
    g_a, _ = run_both(cfg)
    nISM = orc.n_samples(cfg["Tdiff"], cfg["fs"])
    assert 3 * 16 * -(-nISM // 512) >= 256
    e_b = check_parity(g_b, o, nISM)
    e_t = check_parity(g_t, o, nISM)
    assert np.array_equal(g_a, g_b)
    assert e_b < 5e-6 and e_t < 1e-6
    assert rel_l2(g_b[..., :nISM], g_t[..., :nISM]) < 5e-6
    small = W.plain_room()
    g_s, o_s = run_both(small)
    g_st, _ = run_both(small, method="taps")
    g_sb, _ = run_both(small, method="basis")
    assert np.array_equal(g_s, g_st)
    check_parity(g_sb, o_s, orc.n_samples(small["Tdiff"], small["fs"]))


@pytest.mark.parametrize("pattern", ["homni", "card", "hypcard", "subcard", "bidir"])
def test_all_polar_patterns(pattern):
    cfg = W.mid_room(n_src=2, n_rcv=3, seed=3)
    cfg.update(spkr_pattern=pattern, mic_pattern=pattern, Tmax=0.1, Tdiff=0.06)
    cfg["nb_img"] = gpuRIR.t2n(0.06, cfg["room_sz"])
    g, o = run_both(cfg)
    check_parity(g, o, orc.n_samples(0.06, cfg["fs"]))


def test_chunking_and_time_tiles():
    # several image chunks per RIR (partial rows) and several time tiles per RIR
    cfg = W.plain_room()
    g0, o = run_both(cfg)
    g1, _ = run_both(cfg, chunk_images=700, tile_samples=1000)
    check_parity(g1, o, 2800)
    assert rel_l2(g1, g0) < 2e-6


def test_full_ism_no_tail_and_all_tail():
    cfg = W.plain_room()
    cfg["Tmax"] = 0.2
    cfg["Tdiff"] = 0.2
    cfg["nb_img"] = gpuRIR.t2n(0.2, cfg["room_sz"])
    g, o = run_both(cfg)
    check_parity(g, o, o.shape[-1])
    # Tdiff below the direct path: no early part at all; the envelope is degenerate as in the reference
    cfg["Tdiff"] = 0.0
    g, o = run_both(cfg)
    assert np.array_equal(np.isnan(g), np.isnan(o))
    assert np.allclose(np.nan_to_num(g), np.nan_to_num(o), atol=1e-6)


@pytest.mark.parametrize("n_s,n_ism", [(4002, 1602), (4002, 1600), (4004, 1602), (1604, 1602),
                                       (1606, 1602), (1608, 1602), (1854, 1600), (1852, 1600)])
def test_tail_alignment_and_short_tails(n_s, n_ism):
    """Sample counts are even (the reference's rounding), so a row length of 2 mod 4 puts alternate
    receivers' tails 8 bytes off a 16-byte boundary, and an early part of 2 mod 4 does so for all;
    tails of 2, 4 and 6 samples, and of 4 * 63 (+2) samples (a 64-lane range of noise blocks
    advances by 63), cover the ends of the aligned group stores."""
    cfg = W.plain_room()
    rng = np.random.default_rng(5)
    cfg["pos_rcv"] = rng.uniform(0.3, 2.2, size=(5, 3))
    cfg["Tmax"] = (n_s - 0.5) / cfg["fs"]
    cfg["Tdiff"] = (n_ism - 0.5) / cfg["fs"]
    g, o = run_both(cfg)
    assert g.shape == (1, 5, n_s) and o.shape == (1, 5, n_s)
    check_parity(g, o, n_ism)
    # every tail sample is written: the noise is non-zero almost surely
    assert np.count_nonzero(g[..., n_ism:]) == 5 * (n_s - n_ism)


def test_tiny_room_single_image():
    cfg = W.plain_room()
    cfg["nb_img"] = [1, 1, 1]
    g, o = run_both(cfg)
    check_parity(g, o, 2800)


def test_gaussian_noise_and_level():
    cfg = W.plain_room()
    args = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
                nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"])
    g = T.simulate_rir(**args, noise="gaussian", seed=77, call_offset=5, level=0.5).cpu().numpy()
    o = orc.simulate_rir(**args, noise="gaussian", seed=77, call_offset=5, level=0.5)
    check_parity(g, o, 2800)


def test_drop_in_numpy_api_lut_default_and_toggle():
    cfg = W.plain_room()
    args = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
                nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"])
    gpuRIR.activateLUT(True)
    gpuRIR.set_seed(1234, 0)
    g = gpuRIR.simulateRIR(**args)
    assert g.dtype == np.float32 and g.shape == (1, 2, 12800)
    o = orc.simulate_rir(**args, mode=orc.LUT, call_offset=0)
    check_parity(g, o, 2800)
    gpuRIR.activateLUT(False)
    g2 = gpuRIR.simulateRIR(**args)  # second call: call_offset 1
    o2 = orc.simulate_rir(**args, call_offset=1)
    check_parity(g2, o2, 2800)
    gpuRIR.activateLUT(True)


def test_runs_agree():
    cfg = W.mid_room(n_src=2, n_rcv=8)
    a, _ = run_both(cfg)
    b = T.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"], cfg["Tmax"],
                       cfg["fs"], Tdiff=cfg["Tdiff"], spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"],
                       orV_src=cfg["orV_src"], orV_rcv=cfg["orV_rcv"]).cpu().numpy()
    # taps are summed as integers (order-independent) and partial rows in a fixed order: same bits
    assert np.array_equal(a, b.astype(np.float32))
    c = T.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"], cfg["Tmax"],
                       cfg["fs"], Tdiff=cfg["Tdiff"], spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"],
                       orV_src=cfg["orV_src"], orV_rcv=cfg["orV_rcv"], method="basis").cpu().numpy()
    d = T.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"], cfg["nb_img"], cfg["Tmax"],
                       cfg["fs"], Tdiff=cfg["Tdiff"], spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"],
                       orV_src=cfg["orV_src"], orV_rcv=cfg["orV_rcv"], method="basis").cpu().numpy()
    assert np.array_equal(c, d)


@pytest.mark.parametrize("lut", [False, True])
def test_drop_in_and_plan_runs_are_bit_identical(lut):
    """The numpy drop-in with the same noise stream position, and repeated replays of a plan,
    give the same bits."""
    cfg = W.plain_room()
    kw = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
              nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"])
    gpuRIR.activateLUT(lut)
    try:
        gpuRIR.set_seed(1234, 5)
        a = gpuRIR.simulateRIR(**kw)
        gpuRIR.set_seed(1234, 5)
        b = gpuRIR.simulateRIR(**kw)
    finally:
        gpuRIR.activateLUT(True)
    assert np.array_equal(a, b)
    rooms = W.uneven_rooms(n_rooms=16, seed=3)
    plan = T.BatchPlan(rooms, W.FS, seed=7, precision="lut" if lut else "fp32")
    plan.run()
    torch.cuda.synchronize()
    first = plan.out.clone()
    for _ in range(2):
        plan.out.zero_()
        plan.run()
        torch.cuda.synchronize()
        assert torch.equal(plan.out, first)


def test_batch_equals_single_rooms_and_graph_replay():
    rooms = W.uneven_rooms(n_rooms=12, seed=5)
    fs = W.FS
    out, views = T.simulate_batch(rooms, fs, seed=99, call_offset=3)
    torch.cuda.synchronize()
    for k, r in enumerate(rooms):
        single = T.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                                Tdiff=r["Tdiff"], seed=99, call_offset=3, room_index=k)
        assert rel_l2(views[k].cpu().numpy(), single.cpu().numpy()) < 2e-6
        if k < 3:
            o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                                 Tdiff=r["Tdiff"], seed=99, call_offset=3, room_index=k)
            check_parity(views[k].cpu().numpy(), o, orc.n_samples(r["Tdiff"], fs))
    ref = out.clone()
    plan = T.BatchPlan(rooms, fs, seed=99, call_offset=3)
    for rep in range(3):
        plan.out.fill_(float("nan"))
        plan.run()
        torch.cuda.synchronize()
        x = torch.randn(1000, device=plan.out.device).sum().item()  # eager work between replays
        assert np.isfinite(x)
        assert rel_l2(plan.out.cpu().numpy(), ref.cpu().numpy()) < 2e-6, f"replay {rep}"


# ---- mixed precision: packed FP16 windowed sinc, 64-bit fixed-point sums; bar 1e-3 vs FP64
FP16_TOL = 1e-3


def test_mid_room_fp16_packed():
    cfg = W.mid_room()
    g, o = run_both(cfg, precision="fp16")
    e = check_parity(g, o, 3200, tol=FP16_TOL)
    print(f"mid room FP16 rel-L2 {e:.3e}")


@pytest.mark.parametrize("fs", ALL_FS)
def test_fp16_sampling_rates(fs):
    """Mixed precision, both accumulations: the packed-FP16 tap scatter (what the automatic
    choice takes for this small launch) and the basis method with FP16 filter inputs (up to
    64 kHz; above, a basis request is rejected)."""
    cfg = W.plain_room()
    cfg["fs"] = fs
    cfg["Tmax"] = 0.12
    cfg["Tdiff"] = 0.06
    cfg["nb_img"] = gpuRIR.t2n(0.06, cfg["room_sz"])
    g, o = run_both(cfg, precision="fp16")
    e = check_parity(g, o, orc.n_samples(0.06, fs), tol=FP16_TOL)
    print(f"fs {fs:.0f} FP16 taps rel-L2 {e:.3e}")
    if fs <= 64000.0:
        g, o = run_both(cfg, precision="fp16", method="basis")
        e = check_parity(g, o, orc.n_samples(0.06, fs), tol=FP16_TOL)
        print(f"fs {fs:.0f} FP16 basis rel-L2 {e:.3e}")


def test_fp16_basis_is_the_automatic_choice_for_large_launches():
    """The mid room (128 RIRs x 7 tiles) takes the FP16 basis path by default: the same bits as
    an explicit basis request, within 1e-3 of FP64, and within 1e-3 of the FP32 basis result."""
    cfg = W.mid_room()
    g_a, o = run_both(cfg, precision="fp16")
    g_b, _ = run_both(cfg, precision="fp16", method="basis")
    g_f, _ = run_both(cfg, precision="fp32")
    assert np.array_equal(g_a, g_b)
    check_parity(g_a, o, 3200, tol=FP16_TOL)
    assert rel_l2(g_a[..., :3200], g_f[..., :3200]) < FP16_TOL


def test_fp16_drop_in_toggle(capfd):
    cfg = W.plain_room()
    args = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
                nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"])
    gpuRIR.activateLUT(True)
    gpuRIR.activateMixedPrecision(True)
    assert "Disabling the lookup table." in capfd.readouterr().out
    gpuRIR.set_seed(1234, 0)
    g = gpuRIR.simulateRIR(**args)
    o = orc.simulate_rir(**args, call_offset=0)
    check_parity(g, o, 2800, tol=FP16_TOL)
    gpuRIR.activateLUT(True)


def test_fp16_batch_uneven_rooms():
    rooms = W.uneven_rooms(n_rooms=8, seed=11)
    _, views = T.simulate_batch(rooms, W.FS, precision="fp16")
    torch.cuda.synchronize()
    for k, r in enumerate(rooms):
        o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], W.FS,
                             Tdiff=r["Tdiff"], room_index=k)
        check_parity(views[k].cpu().numpy(), o, orc.n_samples(r["Tdiff"], W.FS), tol=FP16_TOL)


def test_persistent_plan_and_autotune():
    """The persistent launch gives the same bits as one workgroup per item, also on graph
    replays (its work counters are reset by the kernel itself); the autotuner re-plans
    to an equivalent result and caches its choice."""
    rooms = W.uneven_rooms(n_rooms=24, seed=13)
    a = T.BatchPlan(rooms, W.FS)
    b = T.BatchPlan(rooms, W.FS, persistent=True)
    ya = a.run().clone()
    for _ in range(3):
        yb = b.run()
        torch.cuda.synchronize()
        assert torch.equal(yb, ya)
    choice = a.autotune()
    assert choice["ms"] > 0
    yc = a.run().clone()
    torch.cuda.synchronize()
    assert rel_l2(yc.cpu().numpy(), ya.cpu().numpy()) < 1e-6
    r = rooms[0]
    o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], W.FS,
                         Tdiff=r["Tdiff"], room_index=0)
    check_parity(a.rirs[0].cpu().numpy(), o, orc.n_samples(r["Tdiff"], W.FS))
    assert T.BatchPlan(rooms, W.FS).autotune() == choice


def test_no_absorption_walls():
    """beta = 1 on every wall: Sabine's T60 is infinite, alpha = 0 and the tail envelope is flat
    (src/gpuRIR_cuda.cu:123-130 divides by a zero absorption area); GPU and oracle agree."""
    cfg = W.plain_room()
    cfg["beta"] = np.ones(6, np.float32)
    cfg["Tmax"] = 0.1
    cfg["Tdiff"] = 0.04
    cfg["nb_img"] = gpuRIR.t2n(0.04, cfg["room_sz"])
    g, o = run_both(cfg)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(o))
    check_parity(g, o, orc.n_samples(0.04, cfg["fs"]))


def test_ragged_batch_several_sources_and_edge_rooms():
    """Rooms with 1-3 sources and 1-5 receivers, one without a diffuse tail, one with a
    one-image set and one whose early part is empty, all in one launch."""
    rng = np.random.default_rng(21)
    rooms = []
    for k in range(7):
        L = rng.uniform([3.0, 3.0, 2.5], [7.0, 6.0, 3.5])
        ns, nr = int(rng.integers(1, 4)), int(rng.integers(1, 6))
        Tmax = float(rng.uniform(0.05, 0.15))
        Tdiff = float(rng.uniform(0.02, Tmax))
        rooms.append(dict(room_sz=L.astype(np.float32), beta=rng.uniform(0.5, 0.95, 6).astype(np.float32),
                          pos_src=(0.5 + rng.uniform(size=(ns, 3)) * (L - 1.0)).astype(np.float32),
                          pos_rcv=(0.5 + rng.uniform(size=(nr, 3)) * (L - 1.0)).astype(np.float32),
                          Tmax=Tmax, Tdiff=Tdiff, nb_img=gpuRIR.t2n(Tdiff, L)))
    rooms[1]["Tdiff"] = rooms[1]["Tmax"]          # no diffuse tail
    rooms[2]["nb_img"] = [1, 1, 1]                # direct path only
    rooms[3]["Tdiff"] = 0.0                       # no early part: all tail (degenerate envelope)
    _, views = T.simulate_batch(rooms, W.FS, seed=5, call_offset=2)
    torch.cuda.synchronize()
    for k, r in enumerate(rooms):
        o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], W.FS,
                             Tdiff=r["Tdiff"], seed=5, call_offset=2, room_index=k)
        g = views[k].cpu().numpy()
        assert g.shape == o.shape
        if k == 3:
            assert np.array_equal(np.isnan(g), np.isnan(o))
            assert np.allclose(np.nan_to_num(g), np.nan_to_num(o), atol=1e-6)
        else:
            check_parity(g, o, orc.n_samples(r["Tdiff"], W.FS))


def test_crossfade_option():
    """Optional linear cross-fade into the diffuse model before Tdiff (0 = the reference's hard
    switch) against the oracle's, on a single room and in a batch."""
    cfg = W.mid_room(n_src=2, n_rcv=3, seed=4)
    args = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
                nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"],
                spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"], orV_src=cfg["orV_src"],
                orV_rcv=cfg["orV_rcv"])
    nISM = orc.n_samples(cfg["Tdiff"], cfg["fs"])
    for X in (1, 160, 1000):
        g = T.simulate_rir(**args, crossfade=X, seed=3, call_offset=1).cpu().numpy()
        o = orc.simulate_rir(**args, crossfade=X, seed=3, call_offset=1)
        check_parity(g, o, nISM)
        assert rel_l2(g[..., nISM - X:nISM], o[..., nISM - X:nISM]) < 1e-5
    rooms = W.uneven_rooms(n_rooms=5, seed=8)
    _, views = T.simulate_batch(rooms, W.FS, crossfade=64)
    torch.cuda.synchronize()
    for k, r in enumerate(rooms):
        o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], W.FS,
                             Tdiff=r["Tdiff"], room_index=k, crossfade=64)
        check_parity(views[k].cpu().numpy(), o, orc.n_samples(r["Tdiff"], W.FS))


def test_basis_tails_independent_of_item_split():
    """Basis method: the RIRs (early part and tail) match the oracle and are bit-identical however
    the early part is split into work items (tiles keep their 512-sample grid and the envelope
    window is summed in a fixed order); a cross-fade matches the oracle; a batch room with no early
    part (Tdiff below every direct path) gives the reference's degenerate envelope."""
    cfg = W.mid_room(n_src=2, n_rcv=8, seed=6)
    nISM = orc.n_samples(cfg["Tdiff"], cfg["fs"])
    g0, o = run_both(cfg, method="basis")
    check_parity(g0, o, nISM)
    for ts in (512, 1536, 4096):
        g1, _ = run_both(cfg, method="basis", tile_samples=ts)
        assert np.array_equal(g1, g0), ts
    args = dict(room_sz=cfg["room_sz"], beta=cfg["beta"], pos_src=cfg["pos_src"], pos_rcv=cfg["pos_rcv"],
                nb_img=cfg["nb_img"], Tmax=cfg["Tmax"], fs=cfg["fs"], Tdiff=cfg["Tdiff"],
                spkr_pattern=cfg["spkr_pattern"], mic_pattern=cfg["mic_pattern"], orV_src=cfg["orV_src"],
                orV_rcv=cfg["orV_rcv"])
    g2 = T.simulate_rir(**args, crossfade=160, method="basis").cpu().numpy()
    o2 = orc.simulate_rir(**args, crossfade=160)
    check_parity(g2, o2, nISM)
    # a batch whose rooms include one with no early part (Tdiff below every direct path) leaves that
    # room's envelope to the finalize kernel
    rooms = W.uneven_rooms(n_rooms=40, seed=9)
    rooms[3]["Tdiff"] = 0.0
    _, views = T.simulate_batch(rooms, W.FS, method="basis")
    torch.cuda.synchronize()
    for k in (0, 3, 39):
        r = rooms[k]
        ob = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], W.FS,
                              Tdiff=r["Tdiff"], room_index=k)
        gb = views[k].cpu().numpy()
        if k == 3:
            assert np.array_equal(np.isnan(gb), np.isnan(ob))
            assert np.allclose(np.nan_to_num(gb), np.nan_to_num(ob), atol=1e-6)
        else:
            check_parity(gb, ob, orc.n_samples(r["Tdiff"], W.FS))


@pytest.mark.parametrize("precision,method,tol", [("fp32", "taps", 1e-5), ("fp32", "basis", 1e-5),
                                                   ("fp16", "taps", 1e-3), ("fp16", "basis", 1e-3)])
def test_source_outside_the_room(precision, method, tol):
    """A source 5 cm beyond a wall and a receiver 2 cm beyond another: images nearer than the
    direct path.  The fixed-point scale is bounded by the nearest image, so no tap is clipped
    (FP32, 64-bit sums) or wraps (mixed precision, 32-bit sums)."""
    L = [4.0, 3.5, 2.7]
    ps = np.array([[4.05, 1.2, 1.1]], np.float32)
    pr = np.array([[2.5, 3.52, 1.3], [1.0, 1.0, 1.0]], np.float32)
    beta = np.full(6, 0.9, np.float32)
    nb = gpuRIR.t2n(0.1, L)
    kw = dict(room_sz=L, beta=beta, pos_src=ps, pos_rcv=pr, nb_img=nb, Tmax=0.14, fs=16000.0, Tdiff=0.1)
    g = T.simulate_rir(**kw, precision=precision, method=method).cpu().numpy()
    o = orc.simulate_rir(**kw)
    check_parity(g, o, orc.n_samples(0.1, 16000.0), tol=tol)


@pytest.mark.parametrize("method", ["taps", "basis"])
def test_delay_fraction_rounding_up_to_one(method):
    """Source (1.5949384, 2.3308206, 1) and receiver (1, 1, 1): the direct path's delay is
    67.99999997 samples, whose fraction (0.99999997) rounds to 1.0 in float32.  The tap at
    j = 1 is then sin(pi (1 - f)) / (pi (1 - f)) with both factors 0 unless the fraction is
    kept below 1; it must come out as the oracle's (sinc ~ 1), and so must the tail, whose
    envelope is fitted on the early part."""
    L = [4.0, 3.5, 2.7]
    ps = np.array([[1.5949384, 2.3308206, 1.0]], np.float32)
    pr = np.array([[1.0, 1.0, 1.0]], np.float32)
    n0, fr, _, _ = orc.images(L, np.full(6, 0.8), ps, pr, [3, 3, 3], 16000.0)
    assert n0[0, 13] == 67 and np.float32(fr[0, 13]) == np.float32(1.0)  # image 13 is the direct path
    kw = dict(room_sz=L, beta=np.full(6, 0.8, np.float32), pos_src=ps, pos_rcv=pr, nb_img=gpuRIR.t2n(0.05, L),
              Tmax=0.08, fs=16000.0, Tdiff=0.05)
    g = T.simulate_rir(**kw, method=method).cpu().numpy()
    check_parity(g, orc.simulate_rir(**kw), orc.n_samples(0.05, 16000.0))


# ---- BF16 filter inputs (the error table's BF16 entry): 8-bit mantissas, so its error is
# stated rather than held to the FP16 mode's 1e-3; bar 1e-2 here
BF16_TOL = 1e-2


@pytest.mark.parametrize("fs", [8000.0, 16000.0, 44100.0, 48000.0])
def test_bf16_basis(fs):
    cfg = W.mid_room(n_src=2, n_rcv=8)
    cfg["fs"] = fs
    cfg["Tmax"] = 0.12
    cfg["Tdiff"] = 0.06
    cfg["nb_img"] = gpuRIR.t2n(0.06, cfg["room_sz"])
    g, o = run_both(cfg, precision="bf16")
    e = check_parity(g, o, orc.n_samples(0.06, fs), tol=BF16_TOL)
    g16, _ = run_both(cfg, precision="fp16", method="basis")
    e16 = check_parity(g16, o, orc.n_samples(0.06, fs), tol=FP16_TOL)
    print(f"fs {fs:.0f}: BF16 rel-L2 {e:.3e}, FP16 {e16:.3e}")
    assert e > e16  # fewer mantissa bits, larger error
    with pytest.raises(ValueError, match="BF16"):
        run_both(cfg, precision="bf16", method="taps")


def test_eager_plan_cache_reuse_and_invalidation():
    """Eager calls reuse the plan of their rooms and options: new positions each call (read from
    device memory), a changed beta (a new plan), and a call on another stream right after the
    call that planned and uploaded it, all against the oracle."""
    base = W.mid_room(n_src=2, n_rcv=8)
    kw = dict(room_sz=base["room_sz"], nb_img=base["nb_img"], Tmax=0.15, fs=base["fs"], Tdiff=0.1,
              spkr_pattern="hypcard", mic_pattern="card")
    for seed in range(3):
        c = W.mid_room(seed=seed, n_src=2, n_rcv=8)
        for beta in (base["beta"], base["beta"] * 0.9):
            args = dict(kw, beta=beta, pos_src=c["pos_src"], pos_rcv=c["pos_rcv"], orV_src=c["orV_src"],
                        orV_rcv=c["orV_rcv"])
            g = T.simulate_rir(**args).cpu().numpy()
            check_parity(g, orc.simulate_rir(**args), orc.n_samples(0.1, kw["fs"]))
    c = W.mid_room(seed=7, n_src=2, n_rcv=8)
    args = dict(kw, beta=base["beta"] * 0.8, pos_src=c["pos_src"], pos_rcv=c["pos_rcv"], orV_src=c["orV_src"],
                orV_rcv=c["orV_rcv"])
    first = T.simulate_rir(**args)  # plans and uploads on the current stream
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        second = T.simulate_rir(**args)  # reuses the plan on another stream
    torch.cuda.synchronize()
    want = orc.simulate_rir(**args)
    check_parity(first.cpu().numpy(), want, orc.n_samples(0.1, kw["fs"]))
    assert torch.equal(first, second)


def test_dense_reflective_room_packed_sums():
    """A 2 x 1.6 x 1.4 m room with beta 0.99 and a 0.25 s early part: ~1.1e6 images per RIR
    (below the 2^21 that selects packed basis sums), up to ~2.5e5 of them in one 512-sample
    tile's distance shell.  The per-tile scale of the packed sums (two basis rows per 64-bit
    LDS word, the narrowest field 24 bits) must leave every field room for the shell's whole
    image count without losing the 1e-5 bar."""
    L = [2.0, 1.6, 1.4]
    ps = np.array([[0.3, 0.4, 0.5]], np.float32)
    pr = np.array([[1.7, 1.2, 0.9], [1.0, 0.8, 0.7]], np.float32)
    nb = gpuRIR.t2n(0.25, L)
    assert 5e5 < np.prod(nb) < 2 ** 21
    kw = dict(room_sz=L, beta=np.full(6, 0.99, np.float32), pos_src=ps, pos_rcv=pr, nb_img=nb, Tmax=0.27,
              fs=16000.0, Tdiff=0.25)
    g = T.simulate_rir(**kw, method="basis").cpu().numpy()
    o = orc.simulate_rir(**kw)
    check_parity(g, o, orc.n_samples(0.25, 16000.0))


def test_more_images_than_packed_sums_take():
    """nb_img with a product above 2^21: the basis kernel keeps 64-bit sums per row (no packing);
    parity with the oracle as for every other room."""
    L = [2.0, 2.0, 2.0]
    ps = np.array([[0.5, 0.6, 0.7]], np.float32)
    pr = np.array([[1.5, 1.3, 1.1]], np.float32)
    nb = gpuRIR.t2n(0.4, L)
    assert np.prod(nb) >= 2 ** 21
    kw = dict(room_sz=L, beta=np.full(6, 0.95, np.float32), pos_src=ps, pos_rcv=pr, nb_img=nb, Tmax=0.42,
              fs=16000.0, Tdiff=0.4)
    g = T.simulate_rir(**kw, method="basis").cpu().numpy()
    o = orc.simulate_rir(**kw)
    check_parity(g, o, orc.n_samples(0.4, 16000.0))
