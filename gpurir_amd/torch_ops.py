"""Device-resident API over the C ABI: torch tensors in and out, work on the
current HIP stream, nothing copied to the host.

  simulate_rir       one room, explicit precision / noise seed / tail level
  image_list         stage (a) per (RIR, image): n0, fraction, amplitude
  simulate_batch     many rooms of uneven sizes in one launch
  BatchPlan          the same, planned once and replayed as a hipGraph
  DiffPlan           a differentiable batch planned once: parameters in device buffers
                     read by the captured graph at every replay (optimisation loops)
  simulate_batch_diff / simulate_rir_diff
                     differentiable versions (torch.autograd) with respect to beta,
                     source and receiver positions and the diffuse-tail level
  trajectory         simulateTrajectory on device tensors (float64 output)
"""
import ctypes
import functools

import numpy as np
import torch

from . import _lib

PATTERNS = {"omni": 0, "homni": 1, "card": 2, "hypcard": 3, "subcard": 4, "bidir": 5}
PRECISIONS = {"fp32": _lib.PREC_FP32, "lut": _lib.PREC_LUT, "fp16": _lib.PREC_FP16, "bf16": _lib.PREC_BF16}
NOISES = {"logistic": _lib.NOISE_LOGISTIC, "gaussian": _lib.NOISE_GAUSSIAN}


def _pat(p):
    return PATTERNS[p] if isinstance(p, str) else int(p)


def _dev_f32(x, dev):
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=dev, dtype=torch.float32).contiguous()
    h = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
    if torch.device(dev).type != "cuda":
        return h.to(dev)
    # through pinned memory without a wait: a copy from pageable memory would wait for the
    # stream's earlier work, so each call would start only after the previous one had ended
    return h.pin_memory().to(device=dev, non_blocking=True)


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


METHODS = {"auto": _lib.METHOD_AUTO, "taps": _lib.METHOD_TAPS, "basis": _lib.METHOD_BASIS}


def options(precision="fp32", noise="logistic", seed=1234, call_offset=0, room_index=0, level=1.0,
            chunk_images=0, tile_samples=0, persistent=False, crossfade=0, method="auto"):
    o = _lib.Options()
    o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.noise = NOISES[noise] if isinstance(noise, str) else int(noise)
    o.seed = int(seed)
    o.call_offset = int(call_offset)
    o.room_index = int(room_index)
    o.level = float(level)
    o.chunk_images = int(chunk_images)
    o.tile_samples = int(tile_samples)
    o.persistent = int(bool(persistent))
    o.crossfade = int(crossfade)
    o.method = METHODS[method] if isinstance(method, str) else int(method)
    return o


def n_samples(T, fs):
    return _lib.load().gpurir_n_samples(float(T), float(fs))


def simulate_rir(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=None, spkr_pattern="omni",
                 mic_pattern="omni", orV_src=None, orV_rcv=None, c=343.0, precision="fp32", noise="logistic",
                 seed=1234, call_offset=0, room_index=0, level=1.0, chunk_images=0, tile_samples=0, out=None,
                 device=None, crossfade=0, method="auto"):
    """simulateRIR on device tensors; returns float32 [n_src, n_rcv, nSamples] on the GPU.

    crossfade: samples of linear cross-fade from the early part into the diffuse model before
    Tdiff (0, the default, is the reference's hard switch).  method: "auto" (the default: "basis"
    when the launch has at least 256 tiles of 512 early samples and fs <= 64 kHz, else "taps"),
    "basis" (per-image sums on a few rows and matrix-core filters: a Chebyshev expansion of the
    taps in the fractional delay, 8 terms in "fp32", in "fp16" the same 8 terms with FP16 rows
    and coefficients at 16 kHz (6 terms and FP16 filter inputs at other rates);
    16 phase rows of the table's 1/16-sample grid in "lut") or "taps" (every image scatters its
    windowed-sinc taps)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ps = _dev_f32(pos_src, dev).reshape(-1, 3)
    pr = _dev_f32(pos_rcv, dev).reshape(-1, 3)
    os_ = torch.zeros_like(ps) if orV_src is None else _dev_f32(orV_src, dev).reshape(-1, 3)
    or_ = torch.zeros_like(pr) if orV_rcv is None else _dev_f32(orV_rcv, dev).reshape(-1, 3)
    if Tdiff is None:
        Tdiff = Tmax
    room = np.ascontiguousarray(np.asarray(room_sz, dtype=np.float32).reshape(3))
    b = np.ascontiguousarray(np.asarray(beta, dtype=np.float32).reshape(6))
    nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(3))
    ns = n_samples(Tmax, fs)
    shape = (ps.shape[0], pr.shape[0], ns)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=dev)
    assert out.shape == shape and out.dtype == torch.float32 and out.is_contiguous()
    o = options(precision, noise, seed, call_offset, room_index, level, chunk_images, tile_samples,
                crossfade=crossfade, method=method)
    _lib.check(_lib.load().gpurir_simulate_rir_device(
        _lib.fptr(room), _lib.fptr(b), ps.data_ptr(), ps.shape[0], pr.data_ptr(), pr.shape[0], os_.data_ptr(),
        or_.data_ptr(), _pat(spkr_pattern), _pat(mic_pattern), _lib.iptr(nb), float(Tdiff), float(Tmax), float(fs),
        float(c), ctypes.byref(o), out.data_ptr(), _stream(dev)))
    return out


def image_list(room_sz, beta, pos_src, pos_rcv, nb_img, fs, spkr_pattern="omni", mic_pattern="omni", orV_src=None,
               orV_rcv=None, c=343.0, device=None):
    """Stage (a): returns (n0 int32, frac float32, amp float32), each [n_src*n_rcv, N] on the GPU."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ps = _dev_f32(pos_src, dev).reshape(-1, 3)
    pr = _dev_f32(pos_rcv, dev).reshape(-1, 3)
    os_ = torch.zeros_like(ps) if orV_src is None else _dev_f32(orV_src, dev).reshape(-1, 3)
    or_ = torch.zeros_like(pr) if orV_rcv is None else _dev_f32(orV_rcv, dev).reshape(-1, 3)
    room = np.ascontiguousarray(np.asarray(room_sz, dtype=np.float32).reshape(3))
    b = np.ascontiguousarray(np.asarray(beta, dtype=np.float32).reshape(6))
    nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(3))
    M = ps.shape[0] * pr.shape[0]
    N = int(np.prod(nb.astype(np.int64)))
    n0 = torch.empty((M, N), dtype=torch.int32, device=dev)
    fr = torch.empty((M, N), dtype=torch.float32, device=dev)
    amp = torch.empty((M, N), dtype=torch.float32, device=dev)
    _lib.check(_lib.load().gpurir_image_list_device(
        _lib.fptr(room), _lib.fptr(b), ps.data_ptr(), ps.shape[0], pr.data_ptr(), pr.shape[0], os_.data_ptr(),
        or_.data_ptr(), _pat(spkr_pattern), _pat(mic_pattern), _lib.iptr(nb), float(fs), float(c), n0.data_ptr(),
        fr.data_ptr(), amp.data_ptr(), _stream(dev)))
    return n0, fr, amp


class _BatchArgs:
    """Host arrays of a batch of rooms plus its device position buffers."""

    def __init__(self, rooms, fs, dev):
        R = len(rooms)
        self.R = R
        self.room_sz = np.ascontiguousarray([r["room_sz"] for r in rooms], dtype=np.float32).reshape(R, 3)
        self.beta = np.ascontiguousarray([r["beta"] for r in rooms], dtype=np.float32).reshape(R, 6)
        self.nb = np.ascontiguousarray([r["nb_img"] for r in rooms], dtype=np.int32).reshape(R, 3)
        ps = [np.asarray(r["pos_src"], dtype=np.float32).reshape(-1, 3) for r in rooms]
        pr = [np.asarray(r["pos_rcv"], dtype=np.float32).reshape(-1, 3) for r in rooms]
        self.n_src = np.ascontiguousarray([p.shape[0] for p in ps], dtype=np.int32)
        self.n_rcv = np.ascontiguousarray([p.shape[0] for p in pr], dtype=np.int32)
        self.Tmax = np.ascontiguousarray([r["Tmax"] for r in rooms], dtype=np.float32)
        self.Tdiff = np.ascontiguousarray([r.get("Tdiff", r["Tmax"]) if r.get("Tdiff") is not None else r["Tmax"]
                                           for r in rooms], dtype=np.float32)
        self.level = np.ascontiguousarray([r.get("level", 1.0) for r in rooms], dtype=np.float32)
        os_ = [np.zeros_like(p) if r.get("orV_src") is None else np.asarray(r["orV_src"], np.float32).reshape(-1, 3)
               for p, r in zip(ps, rooms)]
        or_ = [np.zeros_like(p) if r.get("orV_rcv") is None else np.asarray(r["orV_rcv"], np.float32).reshape(-1, 3)
               for p, r in zip(pr, rooms)]
        self.d_ps = torch.as_tensor(np.concatenate(ps), device=dev)
        self.d_pr = torch.as_tensor(np.concatenate(pr), device=dev)
        self.d_os = torch.as_tensor(np.concatenate(os_), device=dev)
        self.d_or = torch.as_tensor(np.concatenate(or_), device=dev)
        lib = _lib.load()
        self.ns = np.array([lib.gpurir_n_samples(float(t), float(fs)) for t in self.Tmax], dtype=np.int64)
        sizes = self.n_src.astype(np.int64) * self.n_rcv.astype(np.int64) * self.ns
        self.offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]), dtype=np.int64)
        self.total = int(sizes.sum())

    def split(self, out):
        """Views [n_src, n_rcv, nSamples] of each room inside the flat output."""
        res = []
        for k in range(self.R):
            o = int(self.offsets[k])
            n = int(self.n_src[k]) * int(self.n_rcv[k]) * int(self.ns[k])
            res.append(out[o:o + n].view(int(self.n_src[k]), int(self.n_rcv[k]), int(self.ns[k])))
        return res

    def call_args(self):
        ll = ctypes.POINTER(ctypes.c_longlong)
        return (self.R, _lib.fptr(self.room_sz), _lib.fptr(self.beta), _lib.iptr(self.nb), _lib.iptr(self.n_src),
                _lib.iptr(self.n_rcv), _lib.fptr(self.Tdiff), _lib.fptr(self.Tmax), _lib.fptr(self.level),
                self.d_ps.data_ptr(), self.d_pr.data_ptr(), self.d_os.data_ptr(), self.d_or.data_ptr())


def simulate_batch(rooms, fs, spkr_pattern="omni", mic_pattern="omni", c=343.0, precision="fp32",
                   noise="logistic", seed=1234, call_offset=0, room_index=0, chunk_images=0, tile_samples=0,
                   device=None, crossfade=0, method="auto"):
    """rooms: list of dicts with room_sz, beta, pos_src, pos_rcv, nb_img, Tmax [, Tdiff, orV_src, orV_rcv, level].

    Returns (flat float32 output on the GPU, list of per-room [n_src, n_rcv, nSamples] views)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    B = _BatchArgs(rooms, fs, dev)
    out = torch.empty(B.total, dtype=torch.float32, device=dev)
    o = options(precision, noise, seed, call_offset, room_index, chunk_images=chunk_images,
                tile_samples=tile_samples, crossfade=crossfade, method=method)
    _lib.check(_lib.load().gpurir_simulate_batch_device(
        *B.call_args(), _pat(spkr_pattern), _pat(mic_pattern), float(fs), float(c), ctypes.byref(o),
        out.data_ptr(), B.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _stream(dev)))
    return out, B.split(out)


class BatchPlan:
    """A batch of rooms planned once; run() replays the captured hipGraph."""

    def __init__(self, rooms, fs, spkr_pattern="omni", mic_pattern="omni", c=343.0, precision="fp32",
                 noise="logistic", seed=1234, call_offset=0, room_index=0, chunk_images=0, tile_samples=0,
                 persistent=False, device=None, crossfade=0, method="auto"):
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dev = dev
        self.B = _BatchArgs(rooms, fs, dev)
        self.out = torch.empty(self.B.total, dtype=torch.float32, device=dev)
        self.rirs = self.B.split(self.out)
        o = options(precision, noise, seed, call_offset, room_index, chunk_images=chunk_images,
                    tile_samples=tile_samples, persistent=persistent, crossfade=crossfade, method=method)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().gpurir_batch_plan_create(
            *self.B.call_args(), _pat(spkr_pattern), _pat(mic_pattern), float(fs), float(c), ctypes.byref(o),
            self.out.data_ptr(), self.B.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), ctypes.byref(h)))
        self._h = h

    @property
    def n_rirs(self):
        return int(_lib.load().gpurir_batch_plan_n_rirs(self._h))

    def run(self):
        _lib.check(_lib.load().gpurir_batch_plan_run(self._h, _stream(self.dev)))
        return self.out

    def autotune(self):
        """Time candidate work-item shapes and launch forms, keep the fastest (cached per
        process).  Returns dict(chunk_images, tile_samples, persistent, ms)."""
        c, t, p, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        _lib.check(_lib.load().gpurir_batch_plan_autotune(self._h, _stream(self.dev), ctypes.byref(c),
                                                          ctypes.byref(t), ctypes.byref(p), ctypes.byref(ms)))
        return {"chunk_images": c.value, "tile_samples": t.value, "persistent": bool(p.value), "ms": ms.value}

    def count_taps(self):
        """In-window taps (RIR, image, t) of the plan: its algorithmic work."""
        n = ctypes.c_longlong()
        _lib.check(_lib.load().gpurir_batch_plan_count_taps(self._h, ctypes.byref(n), _stream(self.dev)))
        return int(n.value)

    def profile(self):
        """(accumulate ms, finalize ms) of one eager run, from HIP events on the current stream."""
        a, f = ctypes.c_float(), ctypes.c_float()
        _lib.check(_lib.load().gpurir_batch_plan_profile(self._h, _stream(self.dev), ctypes.byref(a), ctypes.byref(f)))
        return float(a.value), float(f.value)

    def adjoint(self, grad_out):
        """Gradients (beta [R,6], pos_src [sum n_src,3], pos_rcv [sum n_rcv,3], level [R]) of
        L given grad_out = dL/d out (flat, the layout of self.out) and the last run()'s output."""
        g = grad_out.detach().to(device=self.dev, dtype=torch.float32).contiguous()
        assert g.numel() == self.out.numel()
        gb = torch.empty((self.B.R, 6), dtype=torch.float32, device=self.dev)
        gs = torch.empty_like(self.B.d_ps)
        gr = torch.empty_like(self.B.d_pr)
        gl = torch.empty(self.B.R, dtype=torch.float32, device=self.dev)
        _lib.check(_lib.load().gpurir_batch_plan_adjoint(self._h, g.data_ptr(), gb.data_ptr(), gs.data_ptr(),
                                                         gr.data_ptr(), gl.data_ptr(), _stream(self.dev)))
        return gb, gs, gr, gl

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().gpurir_batch_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class _PlanFn(torch.autograd.Function):
    """out = simulate(beta, pos_src, pos_rcv, level) through a DiffPlan's captured graph."""

    @staticmethod
    def forward(ctx, beta, pos_src, pos_rcv, level, dp, keep):
        dp._load(beta, pos_src, pos_rcv, level)
        dp.plan.run()
        dp._token += 1
        ctx.dp, ctx.token, ctx.keep = dp, dp._token, keep
        ctx.in_meta = [(t.dtype, t.device) for t in (beta, pos_src, pos_rcv, level)]
        if keep:
            # the call's own output and parameters, so that its backward is right whatever
            # other forward passes run before it
            out = dp.plan.out.clone()
            ctx.saved = (dp.beta.clone(), dp.pos_src.clone(), dp.pos_rcv.clone(), dp.level.clone())
            ctx.save_for_backward(out)
            return out
        return dp.plan.out.detach()

    @staticmethod
    def backward(ctx, grad_out):
        dp = ctx.dp
        if ctx.token != dp._token:
            if not ctx.keep:
                raise RuntimeError("DiffPlan: another forward pass ran before this backward; its output buffer "
                                   "now holds that pass (use keep=True to keep each call's output)")
            (out,) = ctx.saved_tensors
            for dst, src in zip((dp.beta, dp.pos_src, dp.pos_rcv, dp.level), ctx.saved):
                dst.copy_(src)
            dp.plan.out.copy_(out)
            dp._token = ctx.token
        gb, gs, gr, gl = dp.plan.adjoint(grad_out)
        res = [t.to(device=dv, dtype=dt) for t, (dt, dv) in zip((gb, gs, gr, gl), ctx.in_meta)]
        return res[0], res[1], res[2], res[3], None, None


def _diff_desc(rooms):
    """Rooms without their parameters (betas and positions are given as tensors)."""
    desc = []
    for r in rooms:
        ns = int(r["n_src"]) if "n_src" in r else int(np.asarray(r["pos_src"]).reshape(-1, 3).shape[0])
        nr = int(r["n_rcv"]) if "n_rcv" in r else int(np.asarray(r["pos_rcv"]).reshape(-1, 3).shape[0])
        d = dict(r)
        d["beta"] = np.asarray(r["beta"], np.float32) if "beta" in r else np.zeros(6, np.float32)
        d["pos_src"] = np.asarray(r["pos_src"], np.float32) if "pos_src" in r else np.full((ns, 3), 0.5, np.float32)
        d["pos_rcv"] = np.asarray(r["pos_rcv"], np.float32) if "pos_rcv" in r else np.full((nr, 3), 0.5, np.float32)
        desc.append(d)
    return desc


class DiffPlan:
    """A differentiable batch of rooms planned once.

    The betas [R, 6], diffuse levels [R] and source / receiver positions live in device
    buffers that the plan's captured hipGraph reads at every replay
    (gpurir_batch_plan_bind_params), so an optimisation step copies the new parameters
    into them on the device and replays: no host copy, no synchronisation and no
    re-planning per step.  The backward pass is the hand-written adjoint of the plan.

        plan = DiffPlan(rooms, fs)            # rooms: room_sz, nb_img, Tmax [, Tdiff,
                                              # orV_src, orV_rcv, n_src / n_rcv]
        y = plan(beta, pos_src, pos_rcv, level)   # autograd-aware, flat output

    keep=False (the default of __call__) returns the plan's output buffer itself; its
    backward must then run before the next forward pass of the same plan (an
    optimisation loop does exactly that; a violation raises).  keep=True returns a copy
    and restores the call's own output and parameters if its backward comes later.
    The forward pass is the exact FP32 simulation."""

    def __init__(self, rooms, fs, spkr_pattern="omni", mic_pattern="omni", c=343.0, noise="logistic", seed=1234,
                 call_offset=0, room_index=0, chunk_images=0, tile_samples=0, device=None, method="auto",
                 autotune=False):
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dev = dev
        desc = _diff_desc(rooms)
        self.plan = BatchPlan(desc, fs, spkr_pattern, mic_pattern, c, "fp32", noise, seed, call_offset, room_index,
                              chunk_images, tile_samples, device=dev, method=method)
        B = self.plan.B
        self.R = B.R
        self.beta = torch.as_tensor(B.beta, device=dev).clone()
        self.level = torch.as_tensor(B.level, device=dev).clone()
        self.pos_src = B.d_ps  # the position buffers the plan was created with
        self.pos_rcv = B.d_pr
        _lib.check(_lib.load().gpurir_batch_plan_bind_params(self.plan._h, self.beta.data_ptr(),
                                                             self.level.data_ptr()))
        self._token = 0
        self.tune = self.plan.autotune() if autotune else None

    @property
    def out(self):
        return self.plan.out

    def split(self, out):
        return self.plan.B.split(out)

    def _load(self, beta, pos_src, pos_rcv, level):
        assert tuple(beta.shape) == (self.R, 6) and tuple(level.shape) == (self.R,)
        assert tuple(pos_src.shape) == tuple(self.pos_src.shape) and tuple(pos_rcv.shape) == tuple(self.pos_rcv.shape)
        with torch.no_grad():
            self.beta.copy_(beta.detach())
            self.pos_src.copy_(pos_src.detach())
            self.pos_rcv.copy_(pos_rcv.detach())
            self.level.copy_(level.detach())

    def __call__(self, beta, pos_src, pos_rcv, level, keep=False):
        return _PlanFn.apply(beta, pos_src, pos_rcv, level, self, keep)


_DIFF_PLANS = {}  # key -> DiffPlan, the most recent few (simulate_batch_diff)
_DIFF_PLANS_MAX = 2


def simulate_batch_diff(rooms, beta, pos_src, pos_rcv, level, fs, spkr_pattern="omni", mic_pattern="omni",
                        c=343.0, noise="logistic", seed=1234, call_offset=0, room_index=0, chunk_images=0,
                        tile_samples=0, device=None, method="auto"):
    """Differentiable batch simulation.

    rooms: list of dicts with room_sz, nb_img, Tmax [, Tdiff, orV_src, orV_rcv] and the number of
    sources / receivers given by pos_src / pos_rcv rows (keys "n_src", "n_rcv" or the shapes of
    "pos_src" / "pos_rcv").  beta [R,6], pos_src [sum n_src,3], pos_rcv [sum n_rcv,3], level [R] are
    tensors (they may require grad).  The forward pass is the exact FP32 simulation; the backward
    pass is the hand-written adjoint.  method ("auto", "taps" or "basis") selects the
    accumulation of both passes.  The batch is planned once per description (a DiffPlan kept for
    the next calls), so repeated calls with new parameters neither copy them to the host nor
    re-plan.  Returns (flat output, per-room views); every call's output is its own tensor."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    desc = []
    for r in rooms:
        ns = int(r["n_src"]) if "n_src" in r else int(np.asarray(r["pos_src"]).reshape(-1, 3).shape[0])
        nr = int(r["n_rcv"]) if "n_rcv" in r else int(np.asarray(r["pos_rcv"]).reshape(-1, 3).shape[0])
        ov = [None if r.get(k) is None else np.asarray(r[k], np.float32).tobytes() for k in ("orV_src", "orV_rcv")]
        desc.append((tuple(np.asarray(r["room_sz"], np.float32).tolist()), tuple(int(x) for x in r["nb_img"]),
                     float(r["Tmax"]), None if r.get("Tdiff") is None else float(r["Tdiff"]), ns, nr, *ov))
    key = (tuple(desc), float(fs), _pat(spkr_pattern), _pat(mic_pattern), float(c), noise, int(seed),
           int(call_offset), int(room_index), int(chunk_images), int(tile_samples), str(dev), method)
    dp = _DIFF_PLANS.pop(key, None)
    if dp is None:
        dp = DiffPlan(rooms, fs, spkr_pattern, mic_pattern, c, noise, seed, call_offset, room_index, chunk_images,
                      tile_samples, dev, method)
        while len(_DIFF_PLANS) >= _DIFF_PLANS_MAX:
            _DIFF_PLANS.pop(next(iter(_DIFF_PLANS)))
    _DIFF_PLANS[key] = dp
    out = dp(beta, pos_src, pos_rcv, level, keep=True)
    return out, dp.split(out)


def simulate_rir_diff(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=None, level=None,
                      spkr_pattern="omni", mic_pattern="omni", orV_src=None, orV_rcv=None, c=343.0,
                      noise="logistic", seed=1234, call_offset=0, room_index=0, chunk_images=0, tile_samples=0,
                      device=None, method="auto"):
    """Differentiable simulateRIR of one room: beta [6], pos_src [S,3], pos_rcv [R,3] and level
    (scalar tensor or None = 1) may require grad.  Returns [S, R, nSamples] float32 on the GPU."""
    if level is None:
        level = torch.ones((), dtype=torch.float32)
    room = dict(room_sz=room_sz, nb_img=nb_img, Tmax=Tmax, Tdiff=Tdiff, orV_src=orV_src, orV_rcv=orV_rcv,
                n_src=pos_src.shape[0], n_rcv=pos_rcv.shape[0])
    out, views = simulate_batch_diff([room], beta.reshape(1, 6), pos_src, pos_rcv, level.reshape(1), fs,
                                     spkr_pattern, mic_pattern, c, noise, seed, call_offset, room_index,
                                     chunk_images, tile_samples, device, method)
    return views[0]


@functools.lru_cache(maxsize=16)
def _even_windows(n, P):
    """Segment starts of simulateTrajectory without timestamps (gpuRIR/__init__.py:204-208):
    int(p n / P), then n."""
    w = np.ascontiguousarray(np.append((np.arange(P) * (n / P)).astype(int), n), dtype=np.int32)
    w.setflags(write=False)
    return w


def trajectory(source_signal, rirs, timestamps=None, fs=None, device=None):
    """simulateTrajectory on device tensors; returns float64 [n + len - 1, n_rcv] on the GPU."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    sig = _dev_f32(source_signal, dev).reshape(-1)
    r = _dev_f32(rirs, dev)
    n = sig.shape[0]
    P, R, L = r.shape
    if timestamps is None:
        w_ini = _even_windows(n, P)
    else:
        w_ini = np.ascontiguousarray(np.append((np.asarray(timestamps) * fs).astype(int), n), dtype=np.int32)
    out = torch.empty((n + L - 1, R), dtype=torch.float64, device=dev)
    _lib.check(_lib.load().gpurir_trajectory_device(sig.data_ptr(), n, r.data_ptr(), P, R, L, _lib.iptr(w_ini),
                                                    out.data_ptr(), _stream(dev)))
    return out


class TrajectoryPlan:
    """simulateTrajectory bound to device buffers and a segmentation, for repeated calls.

    signal [n] and rirs [P, R, L] are float32 device tensors that the plan keeps and reads at
    every run (refill them in place between runs); out [n + L - 1, R] float64 is the plan's
    result buffer.  With simulateTrajectory's own segmentation (no timestamps) a run launches
    the plan's three kernels and does no other host work; other segmentations run the eager
    path."""

    def __init__(self, signal, rirs, timestamps=None, fs=None):
        if not (isinstance(signal, torch.Tensor) and isinstance(rirs, torch.Tensor)):
            raise TypeError("signal and rirs must be device tensors")
        if not (signal.is_cuda and rirs.is_cuda and signal.dtype == torch.float32 and rirs.dtype == torch.float32
                and signal.is_contiguous() and rirs.is_contiguous() and rirs.dim() == 3 and signal.dim() == 1):
            raise ValueError("signal [n] and rirs [P, R, L] must be contiguous float32 tensors on the GPU")
        self.device = rirs.device
        self.signal, self.rirs = signal, rirs
        n = signal.shape[0]
        P, R, L = rirs.shape
        if timestamps is None:
            w_ini = _even_windows(n, P)
        else:
            w_ini = np.ascontiguousarray(np.append((np.asarray(timestamps) * fs).astype(int), n), dtype=np.int32)
        self.out = torch.empty((n + L - 1, R), dtype=torch.float64, device=self.device)
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gpurir_trajectory_plan_create(signal.data_ptr(), n, rirs.data_ptr(), P, R, L,
                                                                 _lib.iptr(w_ini), self.out.data_ptr(),
                                                                 ctypes.byref(self._h)))

    def run(self):
        """Enqueue one simulateTrajectory of the buffers' contents on the current stream; returns out."""
        _lib.check(_lib.load().gpurir_trajectory_plan_run(self._h, _stream(self.device)))
        return self.out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().gpurir_trajectory_plan_destroy(h)
            except Exception:
                pass
            self._h = None
