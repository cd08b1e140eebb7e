"""The surrogate network's CPU oracle (oracle/mlp.py) and host-side pieces;
no GPU calls."""
import numpy as np

from oracle import mlp as omlp


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 8))
    x[:, :4] = rng.uniform(-1.5, 1.5, (n, 4))
    x[:, 4] = 2.0
    return x


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0.0, 0.2, omlp.NPARAM)
    x, t = _rows(7, 1), rng.uniform(0.0, 0.5, 7)
    loss, g = omlp.loss_grad(p, x, t)
    # a few parameters of every block, central differences
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in omlp.SHAPES])
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(6, b - a), replace=False)
                          for a, b in zip(offs[:-1], offs[1:])])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (omlp.loss_grad(p + e, x, t)[0] - omlp.loss_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(g[k]), (k, fd, g[k])
    assert loss == np.mean((omlp.forward(p, x)[0] - t) ** 2)


def test_oracle_layout_matches_the_library():
    from pyglasstools_amd import engine, surrogate

    assert omlp.NPARAM == 8769 == engine.mlp_nparam()
    p = surrogate.init_params(0)
    assert p.dtype == np.float32 and p.size == omlp.NPARAM
    q = omlp.unpack(p)
    assert q["W1"].shape == (5, 64) and q["W4"].shape == (64,) and not q["b2"].any()


def test_inside_class_rule():
    import torch

    from pyglasstools_amd import surrogate

    x = torch.tensor([[0.1, 0.2, 0.5, 0.5, 2, 0, 0, 0], [0.9, 0.0, 1.2, 0.0, 2, 0, 0, 0],
                      [0.0, 0.99, 0.0, -0.99, 2, 0, 0, 0]], dtype=torch.float32)
    assert surrogate.inside_class(x).tolist() == [True, False, True]
