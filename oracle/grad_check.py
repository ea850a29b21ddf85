"""FP64 central differences of losses built on the C oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's accuracy legs may import this module; the gpurir_amd
package never does.  The function differentiated is the FP64 restatement of
simulateRIR (src/gpuRIR_cuda.cu:245-461, rir_oracle.c) with double-precision
geometry, so a step h = 1e-6 perturbs the parameters exactly.

  central_differences   d loss / d params, (loss(p + h e_i) - loss(p - h e_i)) / 2h
  room_mse_grads        gradients of scale * ||y(beta, src, rcv, level) - target||^2
                        for one room: the loss of the inverse problem restricted to
                        the parameters of that room
"""
import numpy as np

from . import oracle as orc

GROUPS = ("beta", "src", "rcv", "level")


def central_differences(loss, params, h=1e-6):
    """params: dict name -> float64 array (or float); loss(dict) -> float.  Returns a dict of
    arrays of the same shapes (floats stay floats)."""
    base = {k: np.array(v, dtype=np.float64) for k, v in params.items()}
    out = {}
    for name, arr in base.items():
        grad = np.zeros_like(arr)
        for idx in np.ndindex(arr.shape):
            p = {k: v.copy() for k, v in base.items()}
            m = {k: v.copy() for k, v in base.items()}
            p[name][idx] += h
            m[name][idx] -= h
            grad[idx] = (loss(p) - loss(m)) / (2.0 * h)
        out[name] = float(grad) if grad.ndim == 0 else grad
    return out


def room_simulation(room, fs, beta, src, rcv, level, seed=1234, room_index=0, call_offset=0, **kw):
    """The oracle's FP64 simulation of one room with the given parameters."""
    return orc.simulate_rir_f64(room["room_sz"], beta, src, rcv, room["nb_img"], room["Tmax"], fs,
                                Tdiff=room.get("Tdiff"), level=level, seed=seed, room_index=room_index,
                                call_offset=call_offset, **kw)


def room_mse_grads(room, target, scale, fs, seed=1234, room_index=0, call_offset=0, h=1e-6, **kw):
    """FP64 central differences of scale * ||y - target||^2 with respect to the room's six betas,
    its source and receiver positions and the diffuse level, at the room's own parameters."""
    target = np.asarray(target, np.float64)

    def loss(p):
        y = room_simulation(room, fs, p["beta"], p["src"], p["rcv"], float(p["level"]), seed=seed,
                            room_index=room_index, call_offset=call_offset, **kw)
        return float(scale * np.sum((y - target) ** 2))

    params = {"beta": np.asarray(room["beta"], np.float64).reshape(6),
              "src": np.asarray(room["pos_src"], np.float64).reshape(-1, 3),
              "rcv": np.asarray(room["pos_rcv"], np.float64).reshape(-1, 3),
              "level": np.float64(room.get("level", 1.0))}
    return central_differences(loss, params, h)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def group_errors(got, want):
    """Relative L2 error of every gradient group (absolute max where the reference is all zero)."""
    errs = {}
    for k in GROUPS:
        w = np.asarray(want[k], np.float64)
        errs[k] = rel_l2(got[k], w) if np.any(w != 0) else float(np.abs(np.asarray(got[k])).max())
    return errs
