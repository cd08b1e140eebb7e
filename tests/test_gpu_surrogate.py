"""The learned bond-integral surrogate on the GPU, through the C ABI: the
FP32 training kernels and the BF16 matrix-core inference against the FP64
oracle of the network (oracle/mlp.py), the training samples against the
oracle's bond integral, the field form against the plain form, and what
training gains over the untrained network."""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp
from pyglasstools_amd import engine, surrogate, workload

from conftest import scaled_error

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    """Round float32 values to the nearest BF16 (even), kept as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 8), dtype=np.float32)
    x[:, :4] = rng.uniform(-1.5, 1.5, (n, 4))
    x[:, 4] = 2.0
    return x


def _params(seed):
    p = surrogate.init_params(seed).astype(np.float64)
    p += np.random.default_rng(seed + 7).normal(0.0, 0.05, p.size)  # nonzero biases too
    return p.astype(np.float32)


def test_mlp_forward_fp32_matches_oracle(gpu):
    import torch

    p, x = _params(3), _rows(1000, 4)
    dev = torch.device("cuda:0")
    tp, tx = torch.tensor(p, device=dev), torch.tensor(x, device=dev)
    y = torch.empty(1000, dtype=torch.float32, device=dev)
    work = torch.empty(engine.mlp_work_floats(1000), dtype=torch.float32, device=dev)
    engine.mlp_forward(tp, tx, y, work)
    ref = omlp.forward(p, x)[0]
    err = np.abs(y.cpu().numpy() - ref).max()
    print("fp32 forward max abs err %.3e (max |y| %.3f)" % (err, np.abs(ref).max()))
    assert err <= 2e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [257, 5003])
def test_mlp_gradient_matches_oracle(gpu, n):
    """n = 257: one partial 32-sample chunk per weight-gradient workgroup;
    n = 5003: two or three chunks each, the last one ragged."""
    import torch

    p, x = _params(5), _rows(n, 6)
    t = np.random.default_rng(8).uniform(0.0, 0.6, n).astype(np.float32)
    dev = torch.device("cuda:0")
    tp, tx, tt = (torch.tensor(a, device=dev) for a in (p, x, t))
    g = torch.empty_like(tp)
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    work = torch.empty(engine.mlp_work_floats(n), dtype=torch.float32, device=dev)
    engine.mlp_grad(tp, tx, tt, g, loss, work)
    rl, rg = omlp.loss_grad(p, x, t)
    gg = g.cpu().numpy().astype(np.float64)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in omlp.SHAPES])
    for (name, _), a, b in zip(omlp.SHAPES, offs[:-1], offs[1:]):
        rel = np.linalg.norm(gg[a:b] - rg[a:b]) / max(np.linalg.norm(rg[a:b]), 1e-30)
        print("grad %s: rel err %.3e" % (name, rel))
        assert rel <= 1e-4
    assert abs(float(loss.item()) - rl) <= 1e-5 * rl


def test_mlp_infer_bf16_matrix_cores(gpu):
    """With BF16-exact weights and inputs the only differences from the FP64
    network are the BF16 roundings of the hidden activations: a wrong operand
    layout would give errors of order one."""
    import torch

    n = 1000  # not a multiple of the 16 samples per step
    p, x = _bf16_round(_params(9)), _bf16_round(_rows(n, 10))
    dev = torch.device("cuda:0")
    y = torch.empty(n, dtype=torch.float32, device=dev)
    engine.mlp_infer(torch.tensor(p, device=dev), torch.tensor(x, device=dev), y)
    ref = omlp.forward(p, x)[0]
    err = np.abs(y.cpu().numpy() - ref)
    scale = np.abs(ref).max()
    print("bf16 inference: max abs err %.3e, mean %.3e, max |y| %.3f" % (err.max(), err.mean(), scale))
    assert err.max() <= 3e-2 * max(1.0, scale) and err.mean() <= 5e-3 * max(1.0, scale)


def test_bond_samples_match_oracle(gpu):
    import torch

    sysd, grid = workload.config_frame(1, 0)
    d = float(sysd["diam"].max())
    e = engine.Engine(dim=2, diameters=sysd["diam"])
    e.set_system(sysd, grid)
    n = e.bond_samples()
    ref = oracle.ik_compute(sysd, grid, max_rcut=max(1.25 * d, d), fields=("virial",))
    assert n == ref["n_nonzero"]
    dev = torch.device("cuda:0")
    x = torch.empty((n, 8), dtype=torch.float32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    assert e.bond_samples(x, t, n) == n
    xs, ts = x.cpu().numpy().astype(np.float64), t.cpu().numpy()
    assert (xs[:, 4] == 2.0).all() and not xs[:, 5:].any()
    norm = oracle.cg_delta("octic", 2.0, 0.0)  # 15 / (8 pi r_cg^2)
    for k in np.random.default_rng(2).choice(n, 300, replace=False):
        a, b = xs[k, :2] * 2.0, xs[k, 2:4] * 2.0
        w = oracle.bond_fixed("octic", 2.0, 5, (0.0, 0.0, 0.0), (a[0], a[1], 0.0), (b[0] - a[0], b[1] - a[1], 0.0))
        assert abs(ts[k] * norm - w) <= 2e-6 * norm, (k, ts[k] * norm, w)


def test_surrogate_form_without_network_is_the_plain_form(gpu):
    import torch

    sysd, grid = workload.config_frame(1, 1)
    d = float(sysd["diam"].max())
    e = engine.Engine(dim=2, diameters=sysd["diam"])
    e.set_system(sysd, grid)
    plain = e.compute(fields=("virial",))["virial"]
    p = torch.tensor(surrogate.init_params(0), device=torch.device("cuda:0"))
    sur = e.compute_surrogate(p, 1.0, 1.0, -1.0)
    st = e.stats()
    assert st["n_network"] == 0 and st["n_pairs"] == st["n_contrib"] > 0
    assert np.abs(sur - plain).max() <= 1e-13 * np.abs(plain).max()
    ref = oracle.ik_compute(sysd, grid, max_rcut=max(1.25 * d, d), fields=("virial",), with_abs=True)
    assert scaled_error(sur, ref["virial"], ref["virial_abs"]).max() <= 1e-10


def test_trained_network_gain(gpu):
    """Training on the method's own samples takes the network from useless to
    a usable stand-in on pairs whose segment stays inside the ball: the
    held-out error drops by more than 5x against the untrained network (about
    1.1k Adam steps on the samples of two frames; 8x measured).  The class
    bounds are calibrated on a third frame, and the field form is evaluated
    on a fourth, disjoint from both: at a tolerance of the inside bound it
    takes exactly that frame's inside pairs from the network, with a field
    error within the bound's budget.  Below every bound it takes none and the
    field is the plain one (to rounding)."""
    import torch

    dev = torch.device("cuda:0")
    d = 1.62
    e = engine.Engine(dim=2, max_rcut=1.25 * d)
    frames = [workload.config_frame(1, f)[0] for f in range(4)]
    grid = workload.config_frame(1, 0)[1]
    x, t = surrogate.collect_samples(e, frames[:2], grid, dev)
    xh, th = surrogate.collect_samples(e, frames[2:3], grid, dev)
    xe, _ = surrogate.collect_samples(e, frames[3:], grid, dev)
    untrained = torch.tensor(surrogate.init_params(0), device=dev)
    p, hist = surrogate.train(x, t, epochs=12, batch=1024, lr=1e-3, lr_final=1e-5, seed=0)
    assert hist[-1]["loss"] < hist[0]["loss"]
    e_un, e_tr = surrogate.errors(untrained, xh, th), surrogate.errors(p, xh, th)
    ins = surrogate.inside_class(xh)
    print("held-out mean |err|: untrained %.3e, trained %.3e (inside: %.3e)" %
          (e_un.mean().item(), e_tr.mean().item(), e_tr[ins].mean().item()))
    assert e_tr.mean().item() < 0.2 * e_un.mean().item()
    b = surrogate.calibrate(p, xh, th)
    e.set_system(frames[3], grid)
    plain = e.compute(fields=("virial",))["virial"]
    sur = e.compute_surrogate(p, b["inside"], b["crossing"], b["inside"])
    st = e.stats()
    frac = st["n_network"] / st["n_pairs"]
    print("bounds %s; network pairs %.1f%%" % (b, 100 * frac))
    assert st["n_pairs"] == xe.shape[0]
    assert st["n_network"] == int(surrogate.inside_class(xe).sum().item()) > 0
    rel = np.abs(sur - plain).max() / np.abs(plain).max()
    print("surrogate field: max |T - T_plain| / max |T_plain| = %.3e" % rel)
    assert 0.0 < rel < 0.5
    none = e.compute_surrogate(p, b["inside"], b["crossing"], 0.5 * min(b["inside"], b["crossing"]))
    # the pair form sums in pair order, the gather in bond-count order
    assert e.stats()["n_network"] == 0 and np.abs(none - plain).max() <= 1e-13 * np.abs(plain).max()


def test_surrogate_form_rejects_unsupported_settings(gpu):
    """The surrogate stands in for the 2D Octic 5-node rule with FP64 bond
    entries only: 3D systems, the FP32 form and other kernels or orders are
    refused with PGT_ERR_UNSUPPORTED, and a missing frame with PGT_ERR_STATE."""
    import torch

    p = torch.tensor(surrogate.init_params(0), device=torch.device("cuda:0"))
    sysd, grid = workload.config_frame(1, 0)
    e = engine.Engine(dim=2, diameters=sysd["diam"])
    with pytest.raises(engine.PGTError) as ex:
        e.compute_surrogate(p, 0.1, 0.1, 1e-6)
    assert ex.value.code == -3
    s3 = workload.make_system(500, 3, seed=3)
    cases = [(dict(dim=3), s3, workload.cartesian_grid_3d(s3["L"][0], 4)),
             (dict(dim=2, precision="fp32"), sysd, grid),
             (dict(dim=2, cg="rect"), sysd, grid),
             (dict(dim=2, order=6), sysd, grid)]
    for kw, sy, gr in cases:
        e = engine.Engine(diameters=sy["diam"], **kw)
        e.set_system(sy, gr)
        with pytest.raises(engine.PGTError) as ex:
            e.compute_surrogate(p, 0.1, 0.1, 1e-6)
        assert ex.value.code == -5, kw
        with pytest.raises(engine.PGTError):
            e.bond_samples()
