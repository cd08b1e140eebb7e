"""FP64 numpy restatement of the bond-integral surrogate network.

TEST INFRASTRUCTURE ONLY (see the oracle package docstring): the checker for
the HIP training and inference kernels of pyglasstools_amd's surrogate.  The
network has no counterpart in the reference; what it learns is the 5-node
Octic bond integral of FixedPointCGFunc::getBondFunc
(pyglasstools/irvingkirkwood/cgfunc/CoarseGrainFunction.h:85-88), whose
oracle is oracle.bond_fixed.

Layout of the parameter vector (FP32 on the device, FP64 here):
W1[5][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4, tanh after the
three hidden layers, identity output; input rows carry 8 floats of which the
first 5 are used.

Pinned by tests/test_surrogate_cpu.py: the gradient against central finite
differences of the loss, and the parameter count and offsets against the C
header's description.
"""
import numpy as np

N_IN, WIDTH = 5, 64
SHAPES = [("W1", (N_IN, WIDTH)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
          ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def forward(p, x):
    """x: (n, >=5) rows; returns (y (n,), hidden activations)."""
    q = unpack(p)
    x = np.asarray(x, dtype=np.float64)[:, :N_IN]
    h1 = np.tanh(x @ q["W1"] + q["b1"])
    h2 = np.tanh(h1 @ q["W2"] + q["b2"])
    h3 = np.tanh(h2 @ q["W3"] + q["b3"])
    y = h3 @ q["W4"] + q["b4"][0]
    return y, (x, h1, h2, h3)


def loss_grad(p, x, t):
    """Mean-squared loss mean((y - t)^2) and its gradient (NPARAM,)."""
    q = unpack(p)
    y, (x, h1, h2, h3) = forward(p, x)
    n = len(y)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    g = {"W4": h3.T @ dy, "b4": np.array([dy.sum()])}
    dz3 = np.outer(dy, q["W4"]) * (1.0 - h3 * h3)
    g["W3"], g["b3"] = h2.T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ q["W3"].T) * (1.0 - h2 * h2)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ q["W2"].T) * (1.0 - h1 * h1)
    g["W1"], g["b1"] = x.T @ dz1, dz1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])
