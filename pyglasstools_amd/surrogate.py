"""Learned bond-integral surrogate: sampling, training, calibration.

The network (a tanh MLP 5 -> 64 -> 64 -> 64 -> 1, see include/pgt_ik.h) stands
in for the Octic kernel's 5-node bond integral of FixedPointCGFunc::getBondFunc
(pyglasstools/irvingkirkwood/cgfunc/CoarseGrainFunction.h:85-88) in 2D.  It
has no counterpart in the reference.  Everything numeric runs on the GPU in
the engine's own kernels:

* samples: every (grid point, bond) pair of a frame, with its 5-node integral
  as the target (`Engine.bond_samples`);
* training: FP32 forward, mean-squared loss, backward and Adam
  (`pgt_mlp_grad`, `pgt_mlp_adam`), on shuffled minibatches;
* calibration: the largest |network - quadrature| over held-out samples, per
  pair class (both bond ends inside the coarse-graining ball, or not), times a
  safety factor, is the class's error bound -- an empirical bound on a single
  bond integral, not a guarantee and not a tolerance on the field;
* the field form (`Engine.compute_surrogate`) takes the network's value for a
  pair whose class bound is within the tolerance and the quadrature's
  otherwise.

Torch provides device memory and the stream, and its stock kernels do the
bookkeeping around the engine's: the minibatch shuffle and row gather in
training, and the error reductions of calibration.
"""
import time

import numpy as np

from . import engine

ROW = 8  # floats per input row (5 used)


def init_params(seed=0):
    """Xavier-uniform weights, zero biases (FP32 host vector)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out, bias in ((5, 64, 64), (64, 64, 64), (64, 64, 64), (64, 1, 1)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, fan_in * fan_out))
        parts.append(np.zeros(bias))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == engine.mlp_nparam()
    return p


def collect_samples(eng, frames, grid, dev):
    """Input rows and 5-node targets of every pair of the given frames (dicts
    as pyglasstools_amd.workload makes them), concatenated on the device."""
    import torch

    xs, ts = [], []
    for fr in frames:
        eng.set_system(fr, grid)
        n = eng.bond_samples()
        x = torch.empty((n, ROW), dtype=torch.float32, device=dev)
        t = torch.empty((n,), dtype=torch.float32, device=dev)
        eng.bond_samples(x, t, n)
        xs.append(x)
        ts.append(t)
    return torch.cat(xs), torch.cat(ts)


def inside_class(x):
    """True where both bond ends lie inside the ball (|a| < 1 and |b| < 1)."""
    a2 = x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]
    b2 = x[:, 2] * x[:, 2] + x[:, 3] * x[:, 3]
    return (a2 < 1.0) & (b2 < 1.0)


def train(x, t, epochs=20, batch=8192, lr=2e-3, lr_final=1e-4, seed=0, params=None, log=None):
    """Minibatch Adam on (x, t) (device tensors).  The learning rate decays
    geometrically from lr to lr_final.  Returns (params tensor, history)."""
    import torch

    dev = x.device
    n = x.shape[0]
    p = torch.tensor(init_params(seed) if params is None else params, device=dev)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    g = torch.zeros_like(p)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    work = torch.empty(engine.mlp_work_floats(batch), dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed + 1)
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    decay = (lr_final / lr) ** (1.0 / max(1, total - 1))
    hist = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.tensor(rng.permutation(n), device=dev)
        xs, ts = x[perm], t[perm]
        acc = 0.0
        for k in range(steps_per_epoch):
            xb = xs[k * batch:(k + 1) * batch].contiguous()
            tb = ts[k * batch:(k + 1) * batch].contiguous()
            engine.mlp_grad(p, xb, tb, g, loss, work)
            step += 1
            engine.mlp_adam(p, m, v, g, step, float(lr * decay ** (step - 1)))
            if k == steps_per_epoch - 1:
                acc = float(loss.item())
        hist.append({"epoch": ep, "loss": acc, "seconds": time.perf_counter() - t0})
        if log:
            log("surrogate epoch %d: minibatch loss %.3e" % (ep, acc))
    torch.cuda.synchronize()
    return p, hist


def errors(params, x, t):
    """|network (BF16 matrix-core inference) - target| per sample (device)."""
    import torch

    y = torch.empty_like(t)
    engine.mlp_infer(params, x, y)
    return (y - t).abs()


def calibrate(params, x, t, safety=2.0):
    """Error bound per pair class from held-out samples: safety x the largest
    error seen in the class (0 samples: infinity)."""
    e = errors(params, x, t)
    ins = inside_class(x)
    out = {}
    for name, mask in (("inside", ins), ("crossing", ~ins)):
        out[name] = float(e[mask].max().item()) * safety if bool(mask.any()) else float("inf")
    return out
