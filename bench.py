"""Benchmark: RIRs/s of the batched simulation (4096 uneven rooms, one hipGraph replay
per step) on one MI355X, with the plain drop-in rate, accuracy against the FP64 oracle,
the roofline of the dominant kernel and the CPU oracle timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one pass of the simulation over the whole batch of rooms (40,709 RIRs):
image amplitudes/delays, windowed-sinc accumulation, envelope and diffuse tail, with
inputs already resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (packed); non-packed issue caps at half of it
FLOP_PER_TAP = 12               # sub, x2pi, /Tw, cos, +1, x0.5, xpi, sin, /, xamp, xsinc, +acc
FLOP_PER_TAP_BWD = 24           # value and derivative of the windowed sinc, both accumulated
# the FP32 basis kernel on a large 16 kHz batch: 1024-sample tiles, filters from FP16 pairs
BASIS_TILE = 1024
BASIS_MFMA_PER_TILE = 864       # v_mfma_f32_16x16x16_f16 per tile: 3 per 16 K values x 9 K steps x 2 rows x 2 blocks x 8 waves
MFMA_16X16X16_FLOP = 8192
FP16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 matrix (v_mfma_f32_16x16x16_f16 issues at half of it)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary (None if absent)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def read_counters(name):
    """Executed-work utilisation lines (key: value) of a committed rocprofv3 counter summary."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    res = {}
    for line in open(p):
        for key in ("valu_issue_frac", "mfma_busy_frac", "lds_bank_conflict_frac"):
            if line.startswith(key + ":"):
                res[key] = float(line.split(":", 1)[1])
    return res or None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true", help="keep the planner's default work-item shape")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16", "lut"],
                    help="sinc evaluation of the batch measured by --only-batch")
    ap.add_argument("--method", default="auto", choices=["auto", "taps", "basis"],
                    help="accumulation of the batch measured by --only-batch (tap scatter or basis expansion)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rooms", type=int, default=4096)
    ap.add_argument("--cpu-sample-rooms", type=int, default=128)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--only-batch", action="store_true", help="time the batched workload only (for profiler runs)")
    ap.add_argument("--only-gradients", action="store_true",
                    help="time the 256-room forward + adjoint step only (profiler runs)")
    ap.add_argument("--only-trajectory", action="store_true",
                    help="time the trajectory workload only: its RIRs, then convolution + overlap-add (profiler runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (profiler traffic reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus must be 1)")

    import numpy as np
    import torch

    import gpurir_amd as gpuRIR
    from gpurir_amd import torch_ops as T
    from gpurir_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    fs = W.FS

    if args.copy_mb > 0:
        n = args.copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] reference copy: {args.copy_mb} MiB read + {args.copy_mb} MiB written")
        del src, dst

    def measure_trajectory(check=False):
        # ---------------- trajectory: 128 points x 8 mics, full ISM 0.5 s, 10 s of noise
        c5 = W.trajectory()
        rirs5 = T.simulate_rir(c5["room_sz"], c5["beta"], c5["pos_src"], c5["pos_rcv"], c5["nb_img"], c5["Tmax"], fs)
        sig5 = torch.as_tensor(c5["signal"], device=dev)
        for _ in range(3):
            T.trajectory(sig5, rirs5)
        torch.cuda.synchronize()
        # device time of the convolution + overlap-add: a trajectory plan bound to the RIRs and the
        # signal launches its three kernels (events on the stream they run on); the eager call
        # (planning cached, three launches, Python argument handling) timed beside it
        # Each figure is the median of three rounds: these calls are short enough (tens of µs to
        # 1 ms) for a stall of the host to inflate a single round's average several times over.
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def rounds(fn, n, k=3):
            dev_ms, host_ms = [], []
            for _ in range(k):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                e0.record()
                for _ in range(n):
                    fn()
                e1.record()
                torch.cuda.synchronize()
                host_ms.append((time.perf_counter() - t0) * 1e3 / n)
                dev_ms.append(e0.elapsed_time(e1) / n)
            return float(np.median(dev_ms)), float(np.median(host_ms))

        n_tr = 20
        plan5 = T.TrajectoryPlan(sig5, rirs5)
        for _ in range(3):
            plan5.run()
        ms_traj, _ = rounds(plan5.run, n_tr)
        y5 = plan5.run()
        ms_traj_eager, ms_traj_host = rounds(lambda: T.trajectory(sig5, rirs5), n_tr)
        y5e = T.trajectory(sig5, rirs5)
        assert torch.equal(y5, y5e), "trajectory plan and eager call differ"
        P5, R5, L5 = rirs5.shape
        n5 = sig5.numel()
        traj_bytes = 4 * P5 * R5 * L5 + 4 * n5 + 8 * (n5 + L5 - 1) * R5  # algorithmic HBM bytes of stage (d)
        traj_gbs = traj_bytes / (ms_traj / 1e3) / 1e9
        # eager simulation of the 1024 trajectory RIRs (the plan cached by the library, positions
        # from host arrays): 4 warm-up calls (the first ones after lighter work run up to 20%
        # slower), then the median of three rounds of 10
        def traj_rirs():
            T.simulate_rir(c5["room_sz"], c5["beta"], c5["pos_src"], c5["pos_rcv"], c5["nb_img"], c5["Tmax"], fs,
                           out=rirs5)

        for _ in range(4):
            traj_rirs()
        ms_traj_rirs, _ = rounds(traj_rirs, 10)
        log(f"[bench] trajectory: RIRs {ms_traj_rirs:.2f} ms; convolution+OLA {ms_traj:.3f} ms on the device "
            f"({traj_gbs:.0f} GB/s of {traj_bytes / 1e6:.1f} MB algorithmic); eager call {ms_traj_eager:.3f} ms on the "
            f"device clock, {ms_traj_host:.3f} ms on the host")
        res = {"rirs_ms": ms_traj_rirs, "conv_ola_ms": ms_traj, "conv_ola_eager_ms": ms_traj_eager,
               "conv_ola_host_ms": ms_traj_host,
               "conv_ola_gbs": traj_gbs, "conv_ola_bytes": traj_bytes, "conv_ola_frac_hbm": traj_gbs / 8000.0}
        if check:
            # accuracy against the FP64 oracle: the RIRs of two trajectory points, and the whole
            # convolution + float64 overlap-add applied to the GPU's own RIRs; the oracle's overlap-add
            # (numpy FFT in FP64, one thread) is also the CPU baseline of stage (d)
            from oracle import oracle as orc
            g_rirs = rirs5.cpu().numpy()
            pts = [0, P5 - 1]
            o = orc.simulate_rir(c5["room_sz"], c5["beta"], c5["pos_src"][pts], c5["pos_rcv"], c5["nb_img"],
                                 c5["Tmax"], fs)
            res["rirs_rel_l2"] = rel_l2(g_rirs[pts], o)
            t0 = time.perf_counter()
            want = orc.simulate_trajectory(c5["signal"], g_rirs)
            res["cpu_conv_ola_ms"] = (time.perf_counter() - t0) * 1e3
            res["cpu_conv_ola"] = "FP64 numpy FFT overlap-add (the reference's host loop), 1 thread"
            res["conv_ola_rel_l2"] = rel_l2(y5.cpu().numpy(), want)
            log(f"[bench] trajectory accuracy: RIRs {res['rirs_rel_l2']:.2e}, convolution+OLA "
                f"{res['conv_ola_rel_l2']:.2e}; CPU FP64 overlap-add {res['cpu_conv_ola_ms']:.0f} ms")
        del y5, y5e, plan5
        return res

    def measure_gradients(fd_check=True):
        # ---------------- gradients: 256 rooms, MSE against targets simulated from perturbed
        # betas and source positions.  A step is what an optimiser runs: the parameters are
        # copied into the plan's device buffers, the captured forward graph replays, the MSE and
        # dL/dout come from torch autograd, the adjoint kernels run and SGD updates the betas,
        # the sources and the levels on the device.  Nothing goes to the host.
        c4 = W.gradient_rooms(256, seed=0)
        target = T.BatchPlan(W.gradient_targets(c4, seed=1), fs, seed=1234).run().clone()
        dp = T.DiffPlan(c4, fs, seed=1234, autotune=not args.no_autotune)
        beta0 = torch.tensor(np.stack([r["beta"] for r in c4]), device=dev)
        src0 = torch.tensor(np.concatenate([r["pos_src"] for r in c4]), device=dev)
        rcv0 = torch.tensor(np.concatenate([r["pos_rcv"] for r in c4]), device=dev)
        beta = beta0.clone().requires_grad_(True)
        src = src0.clone().requires_grad_(True)
        rcv = rcv0.clone().requires_grad_(True)
        level = torch.ones(len(c4), device=dev, requires_grad=True)
        opt = torch.optim.SGD([beta, src, level], lr=2e4)
        lo = torch.tensor([0.1, 0.1, 0.1], device=dev)
        hi = torch.as_tensor(np.stack([np.asarray(r["room_sz"], np.float32) for r in c4]), device=dev) - 0.1

        def sgd_step():
            opt.zero_grad(set_to_none=True)
            rcv.grad = None
            y = dp(beta, src, rcv, level)
            loss = ((y - target) ** 2).mean()
            loss.backward()
            opt.step()
            with torch.no_grad():
                beta.clamp_(0.0, 1.0)
                src.copy_(torch.minimum(torch.maximum(src, lo), hi))
            return loss

        # FD agreement at the initial parameters, through the same DiffPlan
        y = dp(beta, src, rcv, level)
        ((y - target) ** 2).mean().backward()
        torch.cuda.synchronize()
        fd = None
        if fd_check:
            from oracle import grad_check as gc
            t_host = target.cpu().numpy()
            n_img = [int(np.prod(r["nb_img"])) for r in c4]
            n_ism = [T.n_samples(r["Tdiff"], fs) for r in c4]
            ks = list(dict.fromkeys([0, int(np.argmax(n_img)), int(np.argmax(n_ism)), int(np.argmin(n_img)),
                                     len(c4) - 1]))
            per = T.n_samples(0.5, fs) * 4
            worst = {}
            for k in ks:
                r = c4[k]
                want = gc.room_mse_grads(r, t_host[k * per:(k + 1) * per].reshape(1, 4, -1), 1.0 / y.numel(), fs,
                                         seed=1234, room_index=k)
                got = {"beta": beta.grad[k].double().cpu().numpy(), "src": src.grad[k:k + 1].double().cpu().numpy(),
                       "rcv": rcv.grad[4 * k:4 * k + 4].double().cpu().numpy(), "level": float(level.grad[k])}
                for g, e in gc.group_errors(got, want).items():
                    worst[g] = max(worst.get(g, 0.0), e)
            fd = {"rooms": ks, "rel_l2_max_per_group": worst, "fd_rel_l2_max": max(worst.values()),
                  "h": 1e-6, "reference": "FP64 central differences of the C oracle"}
            log(f"[bench] gradients vs FP64 finite differences on rooms {ks}: max rel-L2 {fd['fd_rel_l2_max']:.2e}")
        for _ in range(max(args.warmup, 2)):
            sgd_step()
        torch.cuda.synchronize()
        n_grad = max(args.steps, 5)
        loss0 = float(((dp(beta.detach(), src.detach(), rcv.detach(), level.detach()) - target) ** 2).mean())
        t0 = time.perf_counter()
        for _ in range(n_grad):
            loss = sgd_step()
        torch.cuda.synchronize()
        ms_grad = (time.perf_counter() - t0) * 1e3 / n_grad
        loss1 = float(loss.detach())
        grads_per_s = len(c4) / (ms_grad / 1e3)
        # replay of the forward graph alone, and device time of the adjoint kernels (events on
        # the current stream, which the adjoint runs on)
        gplan = dp.plan
        t0 = time.perf_counter()
        for _ in range(n_grad):
            gplan.run()
        torch.cuda.synchronize()
        ms_fwd = (time.perf_counter() - t0) * 1e3 / n_grad
        gy = (2.0 / gplan.out.numel()) * (gplan.out - target)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        gplan.adjoint(gy)
        e0.record()
        for _ in range(n_grad):
            gplan.adjoint(gy)
        e1.record()
        torch.cuda.synchronize()
        ms_adj = e0.elapsed_time(e1) / n_grad
        ms_acc_fwd = float(np.mean([gplan.profile()[0] for _ in range(3)]))
        taps4 = gplan.count_taps()
        bwd_tflops = FLOP_PER_TAP_BWD * taps4 / (ms_adj / 1e3) / 1e12
        log(f"[bench] gradients: {ms_grad:.3f} ms per SGD step of {len(c4)} rooms ({grads_per_s:.4e} rooms/s; "
            f"loss {loss0:.3e} -> {loss1:.3e} over {n_grad} steps), forward replay {ms_fwd:.3f} ms, adjoint kernels "
            f"{ms_adj:.3f} ms ({bwd_tflops:.1f} TFLOP/s of 24 flop x {taps4:.3e} taps)")
        return {"grads_per_s": grads_per_s, "ms_per_step": ms_grad, "ms_fwd": ms_fwd, "ms_adjoint": ms_adj,
                "ms_forward_accumulate": ms_acc_fwd, "taps": taps4, "loss_first": loss0, "loss_last": loss1,
                "fd": fd, "forward_autotune": dp.tune,
                "roofline": {"bound": "valu", "achieved": bwd_tflops, "peak": FP32_VALU_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": bwd_tflops / FP32_VALU_PEAK_TFLOPS,
                             "kernel": "adjoint kernels (tail_adjoint, ism_adjoint_basis, grad_rir, grad_param)",
                             "work": f"{taps4} in-window taps x {FLOP_PER_TAP_BWD} flop (value and derivative)"}}

    if args.only_gradients:
        G = measure_gradients(fd_check=False)
        print(json.dumps({"metric": "fwd+bwd grads/s (profiler run)", "value": G["grads_per_s"], "unit": "rooms/s",
                          "ms_per_step": G["ms_per_step"], "ms_fwd": G["ms_fwd"], "ms_adjoint": G["ms_adjoint"]}),
              flush=True)
        return

    if args.only_trajectory:
        tr = measure_trajectory()
        print(json.dumps(dict({"metric": "trajectory (profiler run)"}, **tr)), flush=True)
        return

    # ---------------- batched: 4096 uneven rooms, one graph replay per step
    from oracle import oracle as orc

    log(f"[bench] generating {args.rooms} uneven rooms")
    rooms = W.uneven_rooms(n_rooms=args.rooms, seed=0)

    def measure_batch(precision, method="auto"):
        plan = T.BatchPlan(rooms, fs, precision=precision, seed=1234, call_offset=0, method=method)
        n_rirs = plan.n_rirs
        taps = plan.count_taps()
        tune = plan.autotune() if not args.no_autotune else None
        if tune:
            log(f"[bench] autotuned {precision}: {tune}")
        for _ in range(args.warmup):
            plan.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.run()
        torch.cuda.synchronize()
        ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
        # dominant kernel: accumulation, timed with HIP events on its launch stream
        prof = [plan.profile() for _ in range(5)]
        ms_acc = float(np.mean([p[0] for p in prof]))
        ms_fin = float(np.mean([p[1] for p in prof]))
        # accuracy of the batch against the FP64 oracle (exact sinc, same Philox stream)
        plan.run()
        torch.cuda.synchronize()
        errs, errs_mode = [], []
        for k in range(3):
            r = rooms[k]
            o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                                 Tdiff=r["Tdiff"], seed=1234, call_offset=0, room_index=k)
            g = plan.rirs[k].cpu().numpy()
            errs.append(rel_l2(g, o))
            if precision == "lut":  # and against the oracle's own table mode
                ol = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"],
                                      fs, Tdiff=r["Tdiff"], seed=1234, call_offset=0, room_index=k, mode=orc.LUT)
                errs_mode.append(rel_l2(g, ol))
        res = {"precision": precision, "method": method, "autotune": tune,
               "rirs": n_rirs, "taps": taps, "ms_per_step": ms_step,
               "rirs_per_s": n_rirs / (ms_step / 1e3), "accumulate_ms": ms_acc, "finalize_ms": ms_fin,
               "tflops_algorithmic": FLOP_PER_TAP * taps / (ms_acc / 1e3) / 1e12, "rel_l2_vs_fp64": max(errs),
               "rel_l2_vs_oracle_table_mode": max(errs_mode) if errs_mode else None,
               "out_gb": plan.out.numel() * 4 / 1e9}
        log(f"[bench] batched {precision}: {n_rirs} RIRs, W = {taps:.4e} taps, {ms_step:.3f} ms/step, "
            f"{res['rirs_per_s']:.4e} RIRs/s; accumulate {ms_acc:.3f} ms ({res['tflops_algorithmic']:.2f} TFLOP/s "
            f"algorithmic), finalize {ms_fin:.3f} ms; rel-L2 vs FP64 (3 rooms) {max(errs):.3e}")
        del plan
        torch.cuda.empty_cache()
        return res

    modes = [args.precision] if args.only_batch else ["fp32", "fp32_taps", "fp16", "bf16", "lut"]
    if args.only_batch:
        batch = {args.precision: measure_batch(args.precision, args.method)}
    else:
        batch = {m: (measure_batch("fp32", "taps") if m == "fp32_taps" else measure_batch(m)) for m in modes}
    main_mode = args.precision if args.only_batch else "fp32"
    B = batch[main_mode]
    n_rirs, taps, ms_step, rirs_per_s = B["rirs"], B["taps"], B["ms_per_step"], B["rirs_per_s"]
    ms_acc, ms_fin, achieved, errs = B["accumulate_ms"], B["finalize_ms"], B["tflops_algorithmic"], \
        [B["rel_l2_vs_fp64"]]

    if args.only_batch:
        print(json.dumps({"metric": "batched RIRs/s (profiler run)", "value": rirs_per_s, "unit": "RIRs/s",
                          "ms_per_step": ms_step, "kernel_ms": ms_acc, "taps": taps}), flush=True)
        return

    G = measure_gradients()

    # ---------------- plain drop-in: one room, numpy in / numpy out, FP32 exact sinc
    c1 = W.plain_room()
    gpuRIR.activateLUT(False)
    kw = dict(room_sz=c1["room_sz"], beta=c1["beta"], pos_src=c1["pos_src"], pos_rcv=c1["pos_rcv"],
              nb_img=c1["nb_img"], Tmax=c1["Tmax"], fs=fs, Tdiff=c1["Tdiff"])
    for _ in range(5):
        gpuRIR.simulateRIR(**kw)
    n_plain = 50
    t0 = time.perf_counter()
    for _ in range(n_plain):
        r1 = gpuRIR.simulateRIR(**kw)
    t1 = time.perf_counter()
    plain_rirs_s = 2 * n_plain / (t1 - t0)
    gpuRIR.set_seed(1234, 0)
    r1 = gpuRIR.simulateRIR(**kw)
    err_plain = rel_l2(r1, orc.simulate_rir(**kw, call_offset=0))
    gpuRIR.activateLUT(True)
    log(f"[bench] plain drop-in: {plain_rirs_s:.4e} RIRs/s, rel-L2 {err_plain:.3e}")

    # ---------------- mid room (4 x 32, directional), device resident
    c2 = W.mid_room()
    kw2 = dict(room_sz=c2["room_sz"], beta=c2["beta"], pos_src=c2["pos_src"], pos_rcv=c2["pos_rcv"],
               nb_img=c2["nb_img"], Tmax=c2["Tmax"], fs=fs, Tdiff=c2["Tdiff"], spkr_pattern="hypcard",
               mic_pattern="card", orV_src=c2["orV_src"], orV_rcv=c2["orV_rcv"])
    ps = torch.as_tensor(c2["pos_src"], device=dev)
    pr = torch.as_tensor(c2["pos_rcv"], device=dev)
    kw2d = dict(kw2, pos_src=ps, pos_rcv=pr, orV_src=torch.as_tensor(c2["orV_src"], device=dev),
                orV_rcv=torch.as_tensor(c2["orV_rcv"], device=dev))
    out2 = T.simulate_rir(**kw2d)
    for _ in range(3):
        T.simulate_rir(**kw2d, out=out2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        T.simulate_rir(**kw2d, out=out2)
    torch.cuda.synchronize()
    ms_mid = (time.perf_counter() - t0) * 1e3 / 50
    err_mid = rel_l2(out2.cpu().numpy(), orc.simulate_rir(**kw2))
    log(f"[bench] mid room: {ms_mid:.3f} ms/call ({128 / ms_mid * 1e3:.4e} RIRs/s), rel-L2 {err_mid:.3e}")

    traj = measure_trajectory(check=True)

    # ---------------- the reference's own float32 arithmetic against FP64 (gather order of
    # generateRIR_kernel), early part: the plain room, and 2 x 4 RIRs of the mid room
    fp32_ref = {}
    for name, kwr, sel in (("plain_room", kw, None), ("mid_room_2x4", kw2, (slice(0, 2), slice(0, 4)))):
        k2 = dict(kwr)
        if sel is not None:
            k2["pos_src"], k2["pos_rcv"] = k2["pos_src"][sel[0]], k2["pos_rcv"][sel[1]]
            k2["orV_src"], k2["orV_rcv"] = k2["orV_src"][sel[0]], k2["orV_rcv"][sel[1]]
        n_ism = T.n_samples(k2["Tdiff"], fs)
        ref32 = orc.simulate_rir(**k2, mode=orc.FP32_REF)
        ref64 = orc.simulate_rir(**k2)
        fp32_ref[name] = rel_l2(ref32[..., :n_ism], ref64[..., :n_ism])
    log(f"[bench] the reference's float32 arithmetic against FP64 (early part): {fp32_ref}")

    # ---------------- CPU baseline: the FP64 oracle on a bounded sample of the same batch
    cpu = None
    if not args.skip_cpu:
        nthreads = orc.max_threads()
        model = "unknown CPU"
        try:
            for ln in open("/proc/cpuinfo"):
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
        except OSError:
            pass

        def cpu_run(sample):
            t0 = time.perf_counter()
            for k, r in enumerate(sample):
                orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                                 Tdiff=r["Tdiff"], room_index=k)
            return time.perf_counter() - t0

        sample = rooms[:args.cpu_sample_rooms]
        n_s = sum(len(r["pos_rcv"]) for r in sample)
        cpu_run(sample[:8])  # warm-up
        runs = sorted(cpu_run(sample) for _ in range(3))
        dt = runs[1]
        one = rooms[:max(args.cpu_sample_rooms // 8, 1)]
        n_one = sum(len(r["pos_rcv"]) for r in one)
        orc.set_threads(1)
        dt1 = cpu_run(one)
        orc.set_threads(nthreads)
        cpu = {"value": n_s / dt, "unit": "RIRs/s", "cores": nthreads, "kind": "port",
               "sample": f"first {len(sample)} rooms of the 4096-room batch ({n_s} RIRs), FP64 oracle, windowed "
                         f"scatter, {nthreads} OpenMP threads on {model}: median of 3 runs after a warm-up "
                         f"({', '.join(f'{x:.2f}' for x in runs)} s)",
               "one_thread": {"value": n_one / dt1, "unit": "RIRs/s", "rooms": len(one), "rirs": n_one,
                              "seconds": dt1},
               "cpu_model": model}
        log(f"[bench] CPU oracle: {cpu['value']:.3e} RIRs/s on {nthreads} threads ({model}); "
            f"1 thread {n_one / dt1:.3e} RIRs/s")

    traffic = read_traffic("traffic_ism_basis_uneven_rooms.txt")
    counters = read_counters("counters_ism_basis_uneven_rooms.txt")
    # work the basis kernel executes: its matrix-core filters per 1024-sample tile
    tiles = sum(len(r["pos_src"]) * len(r["pos_rcv"]) * -(-T.n_samples(r["Tdiff"], fs) // BASIS_TILE)
                for r in rooms)
    mfma_tflops = tiles * BASIS_MFMA_PER_TILE * MFMA_16X16X16_FLOP / (ms_acc / 1e3) / 1e12
    line = {
        "metric": "RIRs/s batched vs plain at <=1e-3 rel-L2 vs FP64; fwd+bwd grads/s",
        "value": rirs_per_s,
        "unit": "RIRs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.rooms} uneven rooms in one launch (3-10 x 3-8 x 2.5-4 m, T60 0.2-1.2 s, "
                               "1 src x 4-16 rcv, fs 16 kHz, Tmax 1.1 T60, Tdiff at 15 dB), FP32 exact sinc, "
                               "autotuned work items, hipGraph replay", "rirs": n_rirs, "rooms": args.rooms, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_VALU_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "ism_basis_persistent_kernel" if (B["autotune"] or {}).get("persistent", True)
                     else "ism_basis_kernel", "kernel_ms": ms_acc,
                     "work": f"{taps} in-window taps x {FLOP_PER_TAP} flop (the tap formula's work; the basis "
                             "method executes 8 fixed-point sums per image (4 packed 64-bit LDS adds) and "
                             "matrix-core filters instead)",
                     "executed": {"tiles": tiles, "tile_samples": BASIS_TILE,
                                  "mfma": "v_mfma_f32_16x16x16_f16, FP16 pairs (3 per 16 K values)",
                                  "mfma_per_tile": BASIS_MFMA_PER_TILE, "mfma_tflops": mfma_tflops,
                                  "mfma_frac_of_fp16_dense_peak": mfma_tflops / FP16_MFMA_PEAK_TFLOPS,
                                  "counters": counters}},
        "cpu_baseline": cpu,
        "accuracy": {"batch_rel_l2_max_3_rooms": max(errs), "plain_rel_l2": err_plain, "mid_room_rel_l2": err_mid,
                     "trajectory_rirs_rel_l2": traj.get("rirs_rel_l2"),
                     "trajectory_conv_ola_rel_l2": traj.get("conv_ola_rel_l2"),
                     "gradients_fd_rel_l2_max": (G["fd"] or {}).get("fd_rel_l2_max"),
                     "reference_fp32_arithmetic_rel_l2": fp32_ref},
        "plain_rirs_per_s": plain_rirs_s,
        "batched_over_plain": rirs_per_s / plain_rirs_s,
        "mid_room_ms": ms_mid,
        "finalize_ms": ms_fin,
        "trajectory": traj,
        "batched_modes": batch,
        "gradients": dict({"workload": "256 rooms (4-8 x 3-6 x 2.5-3.5 m, T60 0.3-0.8 s, 1 src x 4 rcv, Tmax "
                                       "0.5 s, Tdiff at 20 dB), MSE to targets from perturbed betas and sources; "
                                       "d/d(6 betas, source xyz, receiver xyz, level); a step is one SGD update of "
                                       "the betas, sources and levels through DiffPlan (device-resident "
                                       "parameters, captured forward graph, hand-written adjoint)"}, **G),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
