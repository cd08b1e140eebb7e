"""The sweep (pgt_ik_sweep_*): the frames of several systems of uneven N (all
2D or all 3D), each with its own box, grid and search radius, computed with
one launch per stage over all of them.  Each system's fields must equal, bit for bit, its
own engine's compute_frames with the tile gather (the kernels' bodies are
the same; only the workgroup-to-frame mapping differs), and the FP64 fields
the oracle at the drop-in tolerance; replays of a captured graph of the
sweep must repeat the eager result; the bond capacity grows as the engine's
does; unsupported forms are refused before any launch."""
import ctypes

import numpy as np
import pytest

import oracle
from pyglasstools_amd import engine, workload

from conftest import scaled_error

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _stack(frames, key, dev):
    import torch

    return torch.tensor(np.stack([f[key] for f in frames]), dtype=torch.float64, device=dev)


def _systems(spec, precision="fp64", rcut_factor=1.25, seed0=7000):
    """spec: (N, frames, grid side) per system; rcut_factor: the pair
    potential's cutoff in units of sigma (the force is zero beyond it), and
    max_rcut that many of the largest diameters."""
    import torch

    dev = torch.device("cuda:0")
    out = []
    for k, (N, FR, NG) in enumerate(spec):
        frames = [workload.make_system(N, 2, seed0 + 100 * k + f) for f in range(FR)]
        L = frames[0]["L"]
        d0 = max(float(f["diam"].max()) for f in frames)
        max_rcut = max(rcut_factor * d0, d0)
        grid = workload.cartesian_grid_2d(L[0], NG)
        e = engine.Engine(dim=2, max_rcut=max_rcut, precision=precision, scaled_rcut=rcut_factor)
        e.set_box(L, [1, 1, 0])
        e.set_grid(grid)
        e.set_gather_mode("tile")
        G = len(grid)
        y = dict(engine=e, nframes=FR, n=N, frames=frames, grid=grid, max_rcut=max_rcut,
                 pos=_stack(frames, "pos", dev), diam=_stack(frames, "diam", dev),
                 mass=_stack(frames, "mass", dev), vel=_stack(frames, "vel", dev))
        for name in ("virial", "kinetic", "ref_virial", "ref_kinetic"):
            y[name] = torch.full((FR, G, 2, 2), float("nan"), dtype=torch.float64, device=dev)
        out.append(y)
    return out


def _engine_reference(systems, fields):
    """Each system through its own engine (compute_frames, tile gather)."""
    nc = 0
    for y in systems:
        y["engine"].compute_frames(y["nframes"], y["n"], y["pos"], y["diam"], y["mass"], y["vel"], None,
                                   fields=fields, d_virial=y["ref_virial"], d_kinetic=y["ref_kinetic"])
        nc += y["engine"].stats()["n_contrib"]
    return nc


def _assert_same(systems, fields, where):
    import torch

    for k, y in enumerate(systems):
        for name in fields:
            assert torch.equal(y[name], y["ref_" + name]), "%s: system %d %s differs from its engine" % (where, k, name)


SPEC = ((700, 3, 20), (2600, 1, 32), (5000, 2, 24), (1200, 9, 17))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sweep_matches_engines_and_oracle(gpu, precision):
    """Four systems (N 700 .. 5000; 1, 2, 3 and 9 frames; grids of 289 to
    1,024 points, none a multiple of 64 but one): the sweep against each
    engine bit for bit, and the FP64 fields against the oracle."""
    fields = ("virial", "kinetic")
    systems = _systems(SPEC, precision)
    nc_engines = _engine_reference(systems, fields)
    sw = engine.Sweep(systems, fields=fields)
    print("sweep %s: run" % precision, flush=True)
    sw.run()
    _assert_same(systems, fields, "sweep " + precision)
    st = sw.stats()
    assert st["n_contrib"] == nc_engines
    assert st["n_particles"] == sum(y["n"] * y["nframes"] for y in systems)
    assert st["n_grid"] == sum(len(y["grid"]) * y["nframes"] for y in systems)
    if precision != "fp64":
        return
    n_nonzero = 0
    for k, y in enumerate(systems):
        for f in range(y["nframes"]):
            ref = oracle.ik_compute(y["frames"][f], y["grid"], fields=fields, max_rcut=y["max_rcut"], with_abs=True)
            n_nonzero += ref["n_nonzero"]
            gv = y["virial"][f].cpu().numpy()
            gk = y["kinetic"][f].cpu().numpy()
            err = scaled_error(gv, ref["virial"], ref["virial_abs"], floor=1e-4).max()
            assert err <= TOL, (k, f, err)
            assert np.abs(gk - ref["kinetic"]).max() <= 1e-12 * np.abs(ref["kinetic"]).max(), (k, f)
    assert st["n_contrib"] == n_nonzero


def test_sweep_graph_replays_repeat_the_run(gpu):
    """The sweep's launches captured into one graph on the sweep's stream and
    replayed three times, with the program's own eager work and NaN-filled
    outputs between replays: every replay equals the eager run, and finish()
    reports the same counts."""
    import torch

    fields = ("virial", "kinetic")
    systems = _systems(SPEC[:3])
    s = torch.cuda.Stream()
    sw = engine.Sweep(systems, fields=fields, stream_handle=s.cuda_stream)
    sw.run()
    ref = {(k, name): y[name].clone() for k, y in enumerate(systems) for name in fields}
    nc = sw.stats()["n_contrib"]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        sw.enqueue()
    sw.finish()  # clears the counters the capture left untouched
    for rep in range(3):
        for y in systems:
            for name in fields:
                y[name].fill_(float("nan"))
        junk = torch.randn(1 << 20, device="cuda:0").sum()  # eager work between replays
        torch.cuda.synchronize()
        print("sweep graph: replay %d" % rep, flush=True)
        g.replay()
        torch.cuda.synchronize()
        sw.finish()
        assert sw.stats()["n_contrib"] == nc
        for k, y in enumerate(systems):
            for name in fields:
                assert torch.equal(y[name], ref[(k, name)]), (rep, k, name)
        del junk
    sw.release()


def test_sweep_mixed_geometries_match_engines_and_oracle(gpu):
    """Systems whose cell geometry differs in kind, in one sweep: a box under
    three cell widths (one cell, minimum image; r_cg 1.0), a box open along y
    (one cell along y, no minimum image there), and a periodic one at cell
    reach 2; each against its engine bit for bit and against the oracle."""
    import torch

    dev = torch.device("cuda:0")
    specs = [(30, 2, 6, (1, 1, 0), 1.0, 1), (800, 3, 16, (1, 0, 0), 2.0, 1), (1500, 2, 20, (1, 1, 0), 2.0, 2)]
    systems = []
    for k, (N, FR, NG, per, rcg, reach) in enumerate(specs):
        frames = [workload.make_system(N, 2, 9000 + 10 * k + f) for f in range(FR)]
        for f in frames:
            f["periodic"] = np.array(per)
        L = frames[0]["L"]
        d0 = max(float(f["diam"].max()) for f in frames)
        max_rcut = max(1.25 * d0, d0)
        grid = workload.cartesian_grid_2d(L[0], NG)
        e = engine.Engine(dim=2, max_rcut=max_rcut, cg_rcut=rcg)
        e.set_box(L, list(per))
        e.set_grid(grid)
        e.set_gather_mode("tile")
        e.set_tuning(cell_reach=reach)
        G = len(grid)
        y = dict(engine=e, nframes=FR, n=N, frames=frames, grid=grid, max_rcut=max_rcut, cg_rcut=rcg,
                 pos=_stack(frames, "pos", dev), diam=_stack(frames, "diam", dev),
                 mass=_stack(frames, "mass", dev), vel=_stack(frames, "vel", dev))
        for name in ("virial", "kinetic", "ref_virial", "ref_kinetic"):
            y[name] = torch.full((FR, G, 2, 2), float("nan"), dtype=torch.float64, device=dev)
        systems.append(y)
    fields = ("virial", "kinetic")
    _engine_reference(systems, fields)
    sw = engine.Sweep(systems, fields=fields)
    sw.run()
    _assert_same(systems, fields, "mixed geometries")
    n_nonzero = 0
    for k, y in enumerate(systems):
        for f in range(y["nframes"]):
            ref = oracle.ik_compute(y["frames"][f], y["grid"], fields=fields, max_rcut=y["max_rcut"],
                                    cg_rcut=y["cg_rcut"], with_abs=True)
            n_nonzero += ref["n_nonzero"]
            err = scaled_error(y["virial"][f].cpu().numpy(), ref["virial"], ref["virial_abs"], floor=1e-4).max()
            print("mixed geometries: system %d frame %d scaled err %.2e" % (k, f, err), flush=True)
            assert err <= TOL, (k, f, err)
            gk = y["kinetic"][f].cpu().numpy()
            assert np.abs(gk - ref["kinetic"]).max() <= 1e-12 * np.abs(ref["kinetic"]).max(), (k, f)
    assert sw.stats()["n_contrib"] == n_nonzero


def test_sweep_grows_bond_capacity(gpu):
    """A pair potential cut off at 2.8 sigma gives about 22 bonds a
    particle, past the initial capacity of 16 of engines that have not run:
    the sweep's run grows the capacity and still matches the engines."""
    fields = ("virial",)
    systems = _systems(((900, 2, 16), (1500, 1, 20)), rcut_factor=2.8)
    sw = engine.Sweep(systems, fields=fields)
    sw.run()
    st = sw.stats()
    print("capacity %d, max bonds %d" % (st["bond_capacity"], st["max_bonds"]), flush=True)
    assert st["max_bonds"] > 16 and st["bond_capacity"] >= st["max_bonds"]
    _engine_reference(systems, fields)
    _assert_same(systems, fields, "capacity")


def test_sweep_kinetic_only(gpu):
    fields = ("kinetic",)
    systems = _systems(SPEC[:2])
    _engine_reference(systems, fields)
    sw = engine.Sweep(systems, fields=fields)
    sw.run()
    _assert_same(systems, fields, "kinetic only")
    assert sw.stats()["n_contrib"] == 0


def _systems_3d(spec, precision="fp64", seed0=7500):
    """spec: (N, frames, grid side) per 3D system."""
    import torch

    dev = torch.device("cuda:0")
    out = []
    for k, (N, FR, NG) in enumerate(spec):
        frames = [workload.make_system(N, 3, seed0 + 100 * k + f) for f in range(FR)]
        L = frames[0]["L"]
        d0 = max(float(f["diam"].max()) for f in frames)
        max_rcut = max(1.25 * d0, d0)
        grid = workload.cartesian_grid_3d(L[0], NG)
        e = engine.Engine(dim=3, max_rcut=max_rcut, precision=precision)
        e.set_box(L, [1, 1, 1])
        e.set_grid(grid)
        e.set_gather_mode("tile")
        G = len(grid)
        y = dict(engine=e, nframes=FR, n=N, frames=frames, grid=grid, max_rcut=max_rcut,
                 pos=_stack(frames, "pos", dev), diam=_stack(frames, "diam", dev),
                 mass=_stack(frames, "mass", dev), vel=_stack(frames, "vel", dev))
        for name in ("virial", "kinetic", "ref_virial", "ref_kinetic"):
            y[name] = torch.full((FR, G, 3, 3), float("nan"), dtype=torch.float64, device=dev)
        out.append(y)
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sweep_3d_matches_engines_and_oracle(gpu, precision):
    """Three 3D systems (N 800 to 3,000; 1 or 2 frames; 6^3 to 8^3 grids): the
    sweep against each engine bit for bit (FP32 records, band pre-tests and
    all), the FP64 fields against the oracle, the FP32 ones within 1e-5 of
    each component's largest |T|."""
    fields = ("virial", "kinetic")
    systems = _systems_3d(((800, 2, 6), (3000, 1, 8), (1700, 2, 7)), precision)
    nc = _engine_reference(systems, fields)
    sw = engine.Sweep(systems, fields=fields)
    print("3d sweep %s: run" % precision, flush=True)
    sw.run()
    _assert_same(systems, fields, "3d sweep " + precision)
    assert sw.stats()["n_contrib"] == nc
    for k, y in enumerate(systems):
        for f in range(y["nframes"]):
            ref = oracle.ik_compute(y["frames"][f], y["grid"], fields=fields, max_rcut=y["max_rcut"], with_abs=True)
            gv = y["virial"][f].cpu().numpy()
            gk = y["kinetic"][f].cpu().numpy()
            if precision == "fp64":
                err = scaled_error(gv, ref["virial"], ref["virial_abs"], floor=1e-4).max()
                assert err <= TOL, (k, f, err)
            else:
                for a in range(3):
                    for b in range(a, 3):
                        r = ref["virial"][:, b, a]
                        assert np.abs(gv[:, b, a] - r).max() <= 1e-5 * np.abs(r).max(), (k, f, a, b)
            assert np.abs(gk - ref["kinetic"]).max() <= 1e-12 * np.abs(ref["kinetic"]).max(), (k, f)


def test_sweep_refuses_unsupported_settings(gpu):
    """Mixed dimensions or precisions, a 2D grid off z = 0, other kernels or
    orders and the elastic field are refused at create; enqueue before a run
    is a state error; particle counts past the tile gather's 2^26 are refused
    before any pointer is read."""
    import torch

    good = _systems(SPEC[:1])
    y = good[0]
    y3 = _systems_3d(((300, 1, 4),))[0]
    with pytest.raises(engine.PGTError) as ex:
        engine.Sweep([y, y3], fields=("virial",))
    assert ex.value.code == -1
    e_off = engine.Engine(dim=2, max_rcut=y["max_rcut"])
    e_off.set_box(y["frames"][0]["L"], [1, 1, 0])
    g_off = y["grid"].copy()
    g_off[:, 2] = 0.5
    e_off.set_grid(g_off)
    cases = [(e_off, -5)]
    for kw in (dict(cg="rect"), dict(order=6)):
        e = engine.Engine(dim=2, max_rcut=y["max_rcut"], **kw)
        e.set_box(y["frames"][0]["L"], [1, 1, 0])
        e.set_grid(y["grid"])
        cases.append((e, -5))
    for e, code in cases:
        with pytest.raises(engine.PGTError) as ex:
            engine.Sweep([dict(y, engine=e)], fields=("virial",))
        assert ex.value.code == code
    with pytest.raises(engine.PGTError) as ex:
        engine.Sweep(good, fields=("virial", "elastic"))
    assert ex.value.code == -5
    e32 = engine.Engine(dim=2, max_rcut=y["max_rcut"], precision="fp32")
    e32.set_box(y["frames"][0]["L"], [1, 1, 0])
    e32.set_grid(y["grid"])
    with pytest.raises(engine.PGTError) as ex:
        engine.Sweep([y, dict(y, engine=e32)], fields=("virial",))
    assert ex.value.code == -1
    sw = engine.Sweep(good, fields=("virial",))
    with pytest.raises(engine.PGTError) as ex:
        sw.enqueue()
    assert ex.value.code == -3
    # the C ABI checks the particle count before it reads any pointer
    a = (engine.SweepSystem * 1)()
    a[0].engine = y["engine"].h.value
    a[0].nframes, a[0].n = 1, (1 << 26) + 1
    a[0].d_pos = a[0].d_diam = a[0].d_out_virial = y["pos"].data_ptr()
    h = ctypes.c_void_p()
    assert engine.lib().pgt_ik_sweep_create(1, a, engine.FIELD_VIRIAL, None, ctypes.byref(h)) == -4
    assert not h.value
    torch.cuda.synchronize()
