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
FP32_MFMA_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix
FLOP_PER_TAP = 12               # sub, x2pi, /Tw, cos, +1, x0.5, xpi, sin, /, xamp, xsinc, +acc
BASIS_TILE = 512                # samples per tile of the basis kernel
BASIS_MFMA_PER_TILE = 576       # v_mfma_f32_16x16x4_f32 per tile at Tw = 128: 8 rows x 36 K steps x 2 blocks
MFMA_16X16X4_FLOP = 2048


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true", help="keep the planner's default work-item shape")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "lut"],
                    help="sinc evaluation of the batch measured by --only-batch")
    ap.add_argument("--method", default="auto", choices=["auto", "taps", "basis"],
                    help="accumulation of the batch measured by --only-batch (tap scatter or basis expansion)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rooms", type=int, default=4096)
    ap.add_argument("--cpu-sample-rooms", type=int, default=400)
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

    def measure_trajectory():
        # ---------------- trajectory: 128 points x 8 mics, full ISM 0.5 s, 10 s of noise
        c5 = W.trajectory()
        rirs5 = T.simulate_rir(c5["room_sz"], c5["beta"], c5["pos_src"], c5["pos_rcv"], c5["nb_img"], c5["Tmax"], fs)
        sig5 = torch.as_tensor(c5["signal"], device=dev)
        for _ in range(3):
            T.trajectory(sig5, rirs5)
        torch.cuda.synchronize()
        # device time of the convolution + overlap-add kernels (events on the stream they run on)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        n_tr = 20
        t0 = time.perf_counter()
        e0.record()
        for _ in range(n_tr):
            y5 = T.trajectory(sig5, rirs5)
        e1.record()
        torch.cuda.synchronize()
        ms_traj_host = (time.perf_counter() - t0) * 1e3 / n_tr
        ms_traj = e0.elapsed_time(e1) / n_tr
        P5, R5, L5 = rirs5.shape
        n5 = sig5.numel()
        traj_bytes = 4 * P5 * R5 * L5 + 4 * n5 + 8 * (n5 + L5 - 1) * R5  # algorithmic HBM bytes of stage (d)
        traj_gbs = traj_bytes / (ms_traj / 1e3) / 1e9
        # eager simulation of the 1024 trajectory RIRs (plan built on the host each call): 4 warm-up
        # calls (the first ones after lighter work run up to 20% slower), then the mean of 10
        for _ in range(4):
            T.simulate_rir(c5["room_sz"], c5["beta"], c5["pos_src"], c5["pos_rcv"], c5["nb_img"], c5["Tmax"], fs,
                           out=rirs5)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(10):
            T.simulate_rir(c5["room_sz"], c5["beta"], c5["pos_src"], c5["pos_rcv"], c5["nb_img"], c5["Tmax"], fs,
                           out=rirs5)
        e1.record()
        torch.cuda.synchronize()
        ms_traj_rirs = e0.elapsed_time(e1) / 10
        log(f"[bench] trajectory: RIRs {ms_traj_rirs:.2f} ms; convolution+OLA {ms_traj:.3f} ms on the device "
            f"({traj_gbs:.0f} GB/s of {traj_bytes / 1e6:.1f} MB algorithmic), {ms_traj_host:.3f} ms per call on the host")
        del y5
        return ms_traj_rirs, ms_traj, ms_traj_host, traj_gbs, traj_bytes

    def measure_gradients():
        # ---------------- gradients: 256 rooms, MSE against targets simulated from perturbed
        # betas and source positions; one step = forward (graph) + dL/dout + adjoint kernels
        c4 = W.gradient_rooms(256, seed=0)
        target = T.BatchPlan(W.gradient_targets(c4, seed=1), fs, seed=1234).run().clone()
        gplan = T.BatchPlan(c4, fs, seed=1234)
        gtune = gplan.autotune() if not args.no_autotune else None

        def grad_step():
            y = gplan.run()
            return gplan.adjoint((2.0 / y.numel()) * (y - target))

        for _ in range(max(args.warmup, 1)):
            grads = grad_step()
        torch.cuda.synchronize()
        n_grad = max(args.steps // 2, 3)
        t0 = time.perf_counter()
        for _ in range(n_grad):
            grads = grad_step()
        torch.cuda.synchronize()
        ms_grad = (time.perf_counter() - t0) * 1e3 / n_grad
        t0 = time.perf_counter()
        for _ in range(n_grad):
            gplan.run()
        torch.cuda.synchronize()
        ms_fwd4 = (time.perf_counter() - t0) * 1e3 / n_grad
        grads_per_s = len(c4) / (ms_grad / 1e3)
        gfinite = all(bool(torch.isfinite(gk).all()) for gk in grads)
        log(f"[bench] gradients: {ms_grad:.3f} ms per fwd+bwd step of {len(c4)} rooms ({grads_per_s:.4e} rooms/s), "
            f"forward alone {ms_fwd4:.3f} ms, finite={gfinite}")
        return ms_grad, ms_fwd4, grads_per_s, gtune

    if args.only_gradients:
        ms_grad, ms_fwd4, grads_per_s, gtune = measure_gradients()
        print(json.dumps({"metric": "fwd+bwd grads/s (profiler run)", "value": grads_per_s, "unit": "rooms/s",
                          "ms_per_fwd_bwd": ms_grad, "ms_fwd": ms_fwd4}), flush=True)
        return

    if args.only_trajectory:
        ms_traj_rirs, ms_traj, ms_traj_host, traj_gbs, traj_bytes = measure_trajectory()
        print(json.dumps({"metric": "trajectory (profiler run)", "rirs_ms": ms_traj_rirs, "conv_ola_ms": ms_traj,
                          "conv_ola_host_ms": ms_traj_host, "conv_ola_gbs": traj_gbs}), flush=True)
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
        errs = []
        for k in range(3):
            r = rooms[k]
            o = orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                                 Tdiff=r["Tdiff"], seed=1234, call_offset=0, room_index=k)
            errs.append(rel_l2(plan.rirs[k].cpu().numpy(), o))
        res = {"precision": precision, "method": method if precision in ("fp32", "fp16") else "taps", "autotune": tune,
               "rirs": n_rirs, "taps": taps, "ms_per_step": ms_step,
               "rirs_per_s": n_rirs / (ms_step / 1e3), "accumulate_ms": ms_acc, "finalize_ms": ms_fin,
               "tflops_algorithmic": FLOP_PER_TAP * taps / (ms_acc / 1e3) / 1e12, "rel_l2_vs_fp64": max(errs),
               "out_gb": plan.out.numel() * 4 / 1e9}
        log(f"[bench] batched {precision}: {n_rirs} RIRs, W = {taps:.4e} taps, {ms_step:.3f} ms/step, "
            f"{res['rirs_per_s']:.4e} RIRs/s; accumulate {ms_acc:.3f} ms ({res['tflops_algorithmic']:.2f} TFLOP/s "
            f"algorithmic), finalize {ms_fin:.3f} ms; rel-L2 vs FP64 (3 rooms) {max(errs):.3e}")
        del plan
        torch.cuda.empty_cache()
        return res

    modes = [args.precision] if args.only_batch else ["fp32", "fp32_taps", "fp16", "lut"]
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

    ms_grad, ms_fwd4, grads_per_s, gtune = measure_gradients()

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

    ms_traj_rirs, ms_traj, ms_traj_host, traj_gbs, traj_bytes = measure_trajectory()

    # ---------------- CPU baseline: the FP64 oracle on a bounded sample of the same batch
    cpu = None
    if not args.skip_cpu:
        nthreads = orc.max_threads()
        sample = rooms[:args.cpu_sample_rooms]
        n_s = sum(len(r["pos_rcv"]) for r in sample)
        t0 = time.perf_counter()
        for k, r in enumerate(sample):
            orc.simulate_rir(r["room_sz"], r["beta"], r["pos_src"], r["pos_rcv"], r["nb_img"], r["Tmax"], fs,
                             Tdiff=r["Tdiff"], room_index=k)
        dt = time.perf_counter() - t0
        cpu = {"value": n_s / dt, "unit": "RIRs/s", "cores": nthreads, "kind": "port",
               "sample": f"first {len(sample)} rooms of the 4096-room batch ({n_s} RIRs), FP64 oracle, "
                         f"windowed scatter, {nthreads} OpenMP threads, {dt:.1f} s"}
        log(f"[bench] CPU oracle: {cpu['value']:.3e} RIRs/s on {nthreads} threads")

    traffic = read_traffic("traffic_ism_basis_uneven_rooms.txt")
    # work the basis kernel executes: its matrix-core filters per 512-sample tile
    tiles = sum(len(r["pos_src"]) * len(r["pos_rcv"]) * -(-T.n_samples(r["Tdiff"], fs) // BASIS_TILE)
                for r in rooms)
    mfma_tflops = tiles * BASIS_MFMA_PER_TILE * MFMA_16X16X4_FLOP / (ms_acc / 1e3) / 1e12
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
                             "method executes 8 fixed-point sums per image and matrix-core filters instead)",
                     "executed": {"tiles": tiles, "mfma_tflops": mfma_tflops,
                                  "mfma_frac": mfma_tflops / FP32_MFMA_PEAK_TFLOPS}},
        "cpu_baseline": cpu,
        "accuracy": {"batch_rel_l2_max_3_rooms": max(errs), "plain_rel_l2": err_plain, "mid_room_rel_l2": err_mid},
        "plain_rirs_per_s": plain_rirs_s,
        "batched_over_plain": rirs_per_s / plain_rirs_s,
        "mid_room_ms": ms_mid,
        "finalize_ms": ms_fin,
        "trajectory": {"rirs_ms": ms_traj_rirs, "conv_ola_ms": ms_traj, "conv_ola_host_ms": ms_traj_host,
                       "conv_ola_gbs": traj_gbs, "conv_ola_bytes": traj_bytes,
                       "conv_ola_frac_hbm": traj_gbs / 8000.0},
        "batched_modes": batch,
        "gradients": {"workload": "256 rooms (4-8 x 3-6 x 2.5-3.5 m, T60 0.3-0.8 s, 1 src x 4 rcv, Tmax 0.5 s, "
                                  "Tdiff at 20 dB), MSE to targets from perturbed betas and sources; "
                                  "d/d(6 betas, source xyz, receiver xyz, level)",
                      "grads_per_s": grads_per_s, "ms_per_fwd_bwd": ms_grad, "ms_fwd": ms_fwd4,
                      "forward_autotune": gtune},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
