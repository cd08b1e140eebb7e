"""The item-balanced tile gather (pgt_ik_set_gather_mode 1) against the oracle
and against the lane gather, through the C ABI: 2D and 3D, FP64 and the FP32
form, virial / kinetic / elastic fields, the fixture trajectory, windows that
wrap a whole periodic axis, 3D tiles whose item lists are processed in
several pieces, a non-periodic axis, cell reach 2, several frames per launch,
and bit-for-bit repeatability."""
import numpy as np
import pytest

import oracle
from pyglasstools_amd import engine, workload
from pyglasstools_amd.gsd_reader import open as gsd_open

from conftest import FIXTURE_GSD, scaled_error

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _engine(dim, max_rcut, precision="fp64", mode="tile", reach=None):
    e = engine.Engine(dim=dim, max_rcut=max_rcut, precision=precision)
    e.set_gather_mode(mode)
    if reach:
        e.set_tuning(cell_reach=reach)
    return e


def _check(got, ref, fields, st=None, tol=TOL):
    for name in fields:
        if name == "kinetic":
            err = np.abs(got[name] - ref[name]).max() / max(np.abs(ref[name]).max(), 1e-300)
        else:
            err = scaled_error(got[name], ref[name], ref["virial_abs"]).max()
        print("  %s: max scaled err %.3e" % (name, err), flush=True)
        assert err <= tol, (name, err)
    if st is not None:
        assert st["n_ball"] == ref["n_ball"]
        if "virial" in fields:
            assert st["n_contrib"] == ref["n_nonzero"]


@pytest.mark.parametrize("N,ng,reach", [(2000, 32, 1), (2000, 32, 2), (500, 40, 1), (9000, 96, 1)])
def test_tile_2d_fp64_virial_kinetic(gpu, N, ng, reach):
    sysd = workload.make_system(N, 2, 8100 + N)
    grid = workload.cartesian_grid_2d(sysd["L"][0], ng)
    mr = max(1.25 * float(sysd["diam"].max()), float(sysd["diam"].max()))
    print("tile 2d N=%d grid %d^2 reach %d" % (N, ng, reach), flush=True)
    e = _engine(2, mr, reach=reach)
    e.set_system(sysd, grid)
    got = e.compute(fields=("virial", "kinetic"))
    ref = oracle.ik_compute(sysd, grid, fields=("virial", "kinetic"), max_rcut=mr, with_abs=True)
    _check(got, ref, ("virial", "kinetic"), e.stats())


def test_tile_fixture_frames_match_oracle(gpu):
    with gsd_open(FIXTURE_GSD) as t:
        for f in (180, 199):
            snap = t[f]
            L = np.array([snap.configuration.box[0], snap.configuration.box[1], 1.0])
            sysd = {"dim": 2, "L": L, "periodic": np.array([1, 1, 0]), "pos": snap.particles.position.astype(np.float64),
                    "diam": snap.particles.diameter.astype(np.float64), "mass": np.ones(snap.particles.N),
                    "vel": np.zeros((snap.particles.N, 3))}
            n = int(L[0] / 0.25)
            p = np.linspace(-L[0] / 2, L[0] / 2, n)
            grid = np.zeros((n * n, 3))
            grid[:, 0] = np.repeat(p, n)
            grid[:, 1] = np.tile(p, n)
            mr = max(1.25 * float(sysd["diam"].max()), float(sysd["diam"].max()))
            print("tile fixture frame %d: N=%d, %d grid points" % (f, snap.particles.N, len(grid)), flush=True)
            e = _engine(2, mr)
            e.set_system(sysd, grid)
            got = e.compute(fields=("virial",))
            ref = oracle.ik_compute(sysd, grid, fields=("virial",), max_rcut=mr, with_abs=True)
            _check(got, ref, ("virial",), e.stats())


def test_tile_elastic_and_fp32_2d(gpu):
    sysd = workload.make_system(3000, 2, 8300)
    rng = np.random.default_rng(5)
    sysd["disp"] = rng.normal(0.0, 0.03, (3000, 3)) * np.array([1, 1, 0])
    grid = workload.cartesian_grid_2d(sysd["L"][0], 40)
    mr = max(1.25 * float(sysd["diam"].max()), float(sysd["diam"].max()))
    print("tile 2d elastic", flush=True)
    e = _engine(2, mr)
    e.set_system(sysd, grid)
    got = e.compute(fields=("virial", "kinetic", "elastic"))
    ref = oracle.ik_compute(sysd, grid, fields=("virial", "kinetic", "elastic"), max_rcut=mr, with_abs=True)
    _check(got, ref, ("virial", "kinetic", "elastic"))
    print("tile 2d fp32 form vs lane fp32 form", flush=True)
    outs = {}
    for mode in ("tile", "lane"):
        e = _engine(2, mr, precision="fp32", mode=mode)
        e.set_system(sysd, grid)
        outs[mode] = (e.compute(fields=("virial",))["virial"], e.stats())
    rel = np.abs(outs["tile"][0] - ref["virial"]).max() / np.abs(ref["virial"]).max()
    print("  fp32 tile err vs oracle %.3e" % rel, flush=True)
    assert rel <= 1e-5
    # both FP32 forms: the same FP32 terms up to the compiler's contraction
    # choices, summed in FP64 in different orders
    assert np.abs(outs["tile"][0] - outs["lane"][0]).max() <= 1e-6 * np.abs(ref["virial"]).max()
    assert outs["tile"][1]["n_contrib"] == outs["lane"][1]["n_contrib"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tile_3d_several_chunks(gpu, precision):
    """A 3D tile lists about 2,100 (point, particle) items, more than its
    768-item list holds, so every tile sorts and integrates its list in
    several pieces."""
    sysd = workload.make_system(6000, 3, 8400)
    grid = workload.cartesian_grid_3d(sysd["L"][0], 16)
    mr = max(1.25 * float(sysd["diam"].max()), float(sysd["diam"].max()))
    print("tile 3d %s N=6000 grid 16^3" % precision, flush=True)
    e = _engine(3, mr, precision=precision)
    e.set_system(sysd, grid)
    got = e.compute(fields=("virial",))
    st = e.stats()
    ref = oracle.ik_compute(sysd, grid, fields=("virial",), max_rcut=mr, with_abs=True)
    if precision == "fp64":
        _check(got, ref, ("virial",), st)
    else:
        worst = 0.0
        for a in range(3):
            for b in range(a, 3):
                worst = max(worst, np.abs(got["virial"][:, b, a] - ref["virial"][:, b, a]).max()
                            / np.abs(ref["virial"][:, b, a]).max())
        print("  fp32 worst component err %.3e" % worst, flush=True)
        assert worst <= 1e-5
        assert st["n_ball"] == ref["n_ball"]


def test_tile_nonperiodic_y_and_frames_per_launch(gpu):
    """A box non-periodic along y (one cell row along y, windows of one row)
    and several frames per launch through compute_frames."""
    import torch

    dev = torch.device("cuda:0")
    frames = [workload.make_system(2500, 2, 8500 + f) for f in range(3)]
    for fr in frames:
        fr["periodic"] = np.array([1, 0, 0])
    L = frames[0]["L"]
    grid = workload.cartesian_grid_2d(L[0], 30)
    mr = max(1.25 * 1.62, 1.62)
    e = _engine(2, mr)
    e.set_box(L, [1, 0, 0])
    e.set_grid(grid)
    e.set_tuning(frames_per_launch=3)
    t = lambda k: torch.tensor(np.stack([f[k] for f in frames]), dtype=torch.float64, device=dev)
    ov = torch.zeros((3, len(grid), 2, 2), dtype=torch.float64, device=dev)
    ok = torch.zeros_like(ov)
    print("tile 2d non-periodic y, 3 frames per launch", flush=True)
    e.compute_frames(3, 2500, t("pos"), t("diam"), t("mass"), t("vel"), None, fields=("virial", "kinetic"),
                     d_virial=ov, d_kinetic=ok)
    torch.cuda.synchronize()
    nn = 0
    for f in range(3):
        ref = oracle.ik_compute(frames[f], grid, fields=("virial", "kinetic"), max_rcut=mr, with_abs=True)
        nn += ref["n_nonzero"]
        _check({"virial": ov[f].cpu().numpy(), "kinetic": ok[f].cpu().numpy()}, ref, ("virial", "kinetic"))
    assert e.stats()["n_contrib"] == nn


def test_tile_matches_lane_on_c2_frames(gpu):
    """Both gathers on 4 frames of the 2D N=65,536 workload (256^2 grid)."""
    import torch

    dev = torch.device("cuda:0")
    F, N = 4, 65536
    frames = [workload.make_system(N, 2, 2000 + f) for f in range(F)]
    L = frames[0]["L"]
    grid = workload.cartesian_grid_2d(L[0], 256)
    mr = max(1.25 * float(frames[0]["diam"].max()), float(frames[0]["diam"].max()))
    t = lambda k: torch.tensor(np.stack([f[k] for f in frames]), dtype=torch.float64, device=dev)
    pos, diam, mass, vel = t("pos"), t("diam"), t("mass"), t("vel")
    res = {}
    for mode in ("tile", "lane"):
        e = _engine(2, mr, mode=mode)
        e.set_box(L, [1, 1, 0])
        e.set_grid(grid)
        e.set_tuning(frames_per_launch=F)
        ov = torch.zeros((F, len(grid), 2, 2), dtype=torch.float64, device=dev)
        ok = torch.zeros_like(ov)
        e.compute_frames(F, N, pos, diam, mass, vel, None, fields=("virial", "kinetic"), d_virial=ov, d_kinetic=ok)
        torch.cuda.synchronize()
        res[mode] = (ov.cpu().numpy(), ok.cpu().numpy(), e.stats())
        print("c2 frames %s: gather %.3f ms" % (mode, res[mode][2]["gather_ms"]), flush=True)
    ref = oracle.ik_compute(frames[0], grid, fields=("virial",), max_rcut=mr, with_abs=True)
    ev = scaled_error(res["tile"][0][0], ref["virial"], ref["virial_abs"]).max()
    print("  tile frame 0 max scaled err %.3e" % ev, flush=True)
    assert ev <= TOL
    scale = np.maximum(np.abs(res["lane"][0]), 1e-6 * np.abs(res["lane"][0]).max())
    assert (np.abs(res["tile"][0] - res["lane"][0]) / scale).max() <= 1e-9
    assert np.abs(res["tile"][1] - res["lane"][1]).max() <= 1e-13 * np.abs(res["lane"][1]).max()
    assert res["tile"][2]["n_contrib"] == res["lane"][2]["n_contrib"]
    assert res["tile"][2]["n_ball"] == res["lane"][2]["n_ball"]


def test_tile_results_repeat_bit_for_bit(gpu):
    """The tile gather's sums have a fixed order: two runs of the same job
    give the same bits (2D FP64 virial + kinetic + elastic, 3D FP32)."""
    import torch

    dev = torch.device("cuda:0")
    for dim, prec, fields, N, ng in ((2, "fp64", ("virial", "kinetic", "elastic"), 9000, 64),
                                     (3, "fp32", ("virial",), 6000, 14)):
        frames = [workload.make_system(N, dim, 8700 + f) for f in range(2)]
        L = frames[0]["L"]
        grid = workload.cartesian_grid_2d(L[0], ng) if dim == 2 else workload.cartesian_grid_3d(L[0], ng)
        e = _engine(dim, max(1.25 * 1.62, 1.62), precision=prec)
        e.set_box(L, [1, 1, 1 if dim == 3 else 0])
        e.set_grid(grid)
        e.set_tuning(frames_per_launch=2)
        t = lambda k: torch.tensor(np.stack([f[k] for f in frames]), dtype=torch.float64, device=dev)
        disp = torch.tensor(np.random.default_rng(3).normal(0, 0.03, (2, N, 3)) * np.array([1, 1, 0 if dim == 2 else 1]),
                            device=dev) if "elastic" in fields else None
        outs = []
        for rep in range(2):
            o = {k: torch.zeros((2, len(grid), dim, dim), dtype=torch.float64, device=dev) for k in fields}
            e.compute_frames(2, N, t("pos"), t("diam"), t("mass"), t("vel"), disp, fields=fields,
                             d_virial=o.get("virial"), d_kinetic=o.get("kinetic"), d_elastic=o.get("elastic"))
            torch.cuda.synchronize()
            outs.append(o)
        for k in fields:
            assert torch.equal(outs[0][k], outs[1][k]), (dim, prec, k)


def test_tile_more_than_255_bonds_runs_the_lane_gather(gpu):
    """A particle with more than 255 bonds does not fit the tile gather's
    one-byte bond count: the tile reports it and the job runs again with the
    lane gather, which still matches the oracle."""
    N, rc = 4096, 10.0
    sysd = workload.make_system(N, 2, 8200)
    grid = workload.cartesian_grid_2d(sysd["L"][0], 16)
    d = float(sysd["diam"].max())
    mr = rc * d
    params = oracle.poly12_params(rcut=rc)
    print("tile 2d N=%d with pair cutoff %.1f sigma" % (N, rc), flush=True)
    e = engine.Engine(dim=2, params=params, scaled_rcut=rc, max_rcut=mr)
    e.set_gather_mode("tile")
    e.set_system(sysd, grid)
    got = e.compute(fields=("virial",))
    st = e.stats()
    assert st["max_bonds"] > 255
    ref = oracle.ik_compute(sysd, grid, params=params, scaled_rcut=rc, fields=("virial",), max_rcut=mr,
                            with_abs=True)
    _check(got, ref, ("virial",), st)


def test_tile_irregular_grid_dilute_system_wide_ball(gpu):
    """Edge cases of the tile gather at once: a grid of 37 x 41 points (tiles
    not filled to 64), shifted so that a third of it lies outside the box
    (minimum image and wrapped cells), particles removed from half of the box
    (empty cells, tiles with no items), and r_cg = 3.5, which puts about
    40 particles in a ball and makes every busy tile flush its list several
    times."""
    sysd = workload.make_system(3000, 2, 8600)
    keep = sysd["pos"][:, 0] < 0.0
    sysd = dict(sysd, pos=sysd["pos"][keep], diam=sysd["diam"][keep], mass=sysd["mass"][keep],
                vel=sysd["vel"][keep])
    L = sysd["L"]
    gx = np.linspace(-L[0] / 2, L[0] / 2, 37) + L[0] / 3
    gy = np.linspace(-L[1] / 2, L[1] / 2, 41)
    grid = np.zeros((37 * 41, 3))
    grid[:, 0] = np.repeat(gx, 41)
    grid[:, 1] = np.tile(gy, 37)
    mr = max(1.25 * 1.62, 1.62)
    print("tile 2d dilute N=%d, 37x41 shifted grid, r_cg 3.5" % len(sysd["diam"]), flush=True)
    e = engine.Engine(dim=2, max_rcut=mr, cg_rcut=3.5)
    e.set_gather_mode("tile")
    e.set_system(sysd, grid)
    got = e.compute(fields=("virial", "kinetic"))
    ref = oracle.ik_compute(sysd, grid, fields=("virial", "kinetic"), max_rcut=mr, cg_rcut=3.5, with_abs=True)
    assert (np.abs(ref["virial"]).reshape(len(grid), -1).max(axis=1) == 0).any()  # some empty balls
    _check(got, ref, ("virial", "kinetic"), e.stats())


def test_tile_3d_elastic_and_kinetic(gpu):
    """3D virial, kinetic and elastic fields.  With the elastic field a bond
    is listed when either the force at the separation or at the displaced
    separation is nonzero, so the contribution count is compared with the
    lane gather's (the oracle counts nonzero-force triples)."""
    sysd = workload.make_system(2000, 3, 8700)
    rng = np.random.default_rng(6)
    sysd["disp"] = rng.normal(0.0, 0.03, (2000, 3))
    grid = workload.cartesian_grid_3d(sysd["L"][0], 9)
    mr = max(1.25 * float(sysd["diam"].max()), float(sysd["diam"].max()))
    print("tile 3d elastic + kinetic N=2000 grid 9^3", flush=True)
    ref = oracle.ik_compute(sysd, grid, fields=("virial", "kinetic", "elastic"), max_rcut=mr, with_abs=True)
    stats = {}
    for mode in ("tile", "lane"):
        e = _engine(3, mr, mode=mode)
        e.set_system(sysd, grid)
        got = e.compute(fields=("virial", "kinetic", "elastic"))
        _check(got, ref, ("virial", "kinetic", "elastic"))
        stats[mode] = e.stats()
    assert stats["tile"]["n_ball"] == stats["lane"]["n_ball"] == ref["n_ball"]
    assert stats["tile"]["n_contrib"] == stats["lane"]["n_contrib"]


def test_tile_tiny_system_single_cell_axes(gpu):
    """Twenty particles in a box a little over twice the search radius: one
    cell along each axis, so every window is the whole box."""
    sysd = workload.make_system(20, 2, 8800)
    grid = workload.cartesian_grid_2d(sysd["L"][0], 7)
    mr = max(1.25 * float(sysd["diam"].max()), float(sysd["diam"].max()))
    assert 2 * mr < sysd["L"][0]
    print("tile 2d N=20 L=%.2f" % sysd["L"][0], flush=True)
    e = _engine(2, mr)
    e.set_system(sysd, grid)
    got = e.compute(fields=("virial", "kinetic"))
    ref = oracle.ik_compute(sysd, grid, fields=("virial", "kinetic"), max_rcut=mr, with_abs=True)
    _check(got, ref, ("virial", "kinetic"), e.stats())
