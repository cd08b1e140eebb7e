"""pgt_ik_compute_frames_noalloc (the entry point that allocates nothing: its
work arrays live in a static device scratch array) against
pgt_ik_compute_frames on the same inputs, through the C ABI.  Two calls in a
row, the second with fewer frames and fewer particles (the scratch array
keeps the first call's contents, so anything left uncleared would show), for
the FP64 virial + kinetic form, the elastic field and the 3D FP32 form, with
the lane and the tile gather; and a
frame too large for the scratch array is refused before any launch."""
import numpy as np
import pytest

from pyglasstools_amd import engine, workload

pytestmark = pytest.mark.gpu


def _frames(N, dim, nf, seed0):
    return [workload.make_system(N, dim, seed0 + f) for f in range(nf)]


def _stack(frames, key, dev):
    import torch

    return torch.tensor(np.stack([f[key] for f in frames]), dtype=torch.float64, device=dev)


@pytest.mark.parametrize("gather", ["lane", "tile"])
@pytest.mark.parametrize("dim,precision,fields,ng", [
    (2, "fp64", ("virial", "kinetic"), 32),
    (2, "fp64", ("virial", "kinetic", "elastic"), 24),
    (3, "fp32", ("virial",), 12),
])
def test_noalloc_matches_compute_frames(gpu, dim, precision, fields, ng, gather):
    import torch

    dev = torch.device("cuda:0")
    e = engine.Engine(dim=dim, max_rcut=max(1.25 * 1.62, 1.62), precision=precision)
    e.set_gather_mode(gather)
    for call, (N, nf, seed0) in enumerate(((3000, 3, 7100), (1500, 2, 7200))):
        print("noalloc: %s gather, dim %d %s %s call %d N=%d frames=%d" % (gather, dim, precision, fields, call, N,
                                                                          nf), flush=True)
        frames = _frames(N, dim, nf, seed0)
        L = frames[0]["L"]
        grid = workload.cartesian_grid_2d(L[0], ng) if dim == 2 else workload.cartesian_grid_3d(L[0], ng)
        G = len(grid)
        e.set_box(L, [1, 1, 1 if dim == 3 else 0])
        e.set_grid(grid)
        pos, diam = _stack(frames, "pos", dev), _stack(frames, "diam", dev)
        mass, vel = _stack(frames, "mass", dev), _stack(frames, "vel", dev)
        disp = None
        if "elastic" in fields:
            rng = np.random.default_rng(seed0)
            disp = torch.tensor(rng.normal(0.0, 0.02, (nf, N, 3)) * np.array([1, 1, 0]), device=dev)
        outs = {}
        for mode in ("frames", "noalloc"):
            o = {k: torch.full((nf, G, dim, dim), np.nan, dtype=torch.float64, device=dev) for k in fields}
            fn = e.compute_frames if mode == "frames" else e.compute_frames_noalloc
            fn(nf, N, pos, diam, mass, vel, disp, fields=fields, d_virial=o.get("virial"),
               d_kinetic=o.get("kinetic"), d_elastic=o.get("elastic"))
            torch.cuda.synchronize()
            outs[mode] = ({k: v.cpu().numpy() for k, v in o.items()}, e.stats())
        (a, sa), (b, sb) = outs["frames"], outs["noalloc"]
        for k in fields:
            assert np.isfinite(b[k]).all(), k
            scale = max(np.abs(a[k]).max(), 1e-300)
            err = np.abs(a[k] - b[k]).max() / scale
            print("noalloc: %s max |diff| / max %.3e" % (k, err), flush=True)
            assert err <= 1e-12, (k, err)
        if "virial" in fields or "elastic" in fields:
            assert sa["n_contrib"] == sb["n_contrib"]
        assert sa["n_ball"] == sb["n_ball"]


def test_noalloc_refuses_a_frame_larger_than_its_scratch(gpu):
    import torch

    dev = torch.device("cuda:0")
    N = 300_000  # ~450 bytes of work memory per 2D particle: about 130 MiB for one frame
    e = engine.Engine(dim=2, max_rcut=max(1.25 * 1.62, 1.62))
    L = np.sqrt(N)
    e.set_box([L, L, 1.0], [1, 1, 0])
    e.set_grid(workload.cartesian_grid_2d(L, 8))
    pos = torch.zeros((1, N, 3), dtype=torch.float64, device=dev)
    diam = torch.ones((1, N), dtype=torch.float64, device=dev)
    out = torch.zeros((1, 64, 2, 2), dtype=torch.float64, device=dev)
    with pytest.raises(engine.PGTError, match="PGT_ERR_RANGE"):
        e.compute_frames_noalloc(1, N, pos, diam, fields=("virial",), d_virial=out)
