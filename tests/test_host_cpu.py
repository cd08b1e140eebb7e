"""Host-side logic, without a GPU: GSD ingest, grids, the pybind surface,
.xyz output format, and the frame loop's checkpoint / resume / interrupt
behaviour (pyglasstools/__init__.py:107-138, utils.py:71-83)."""
import os
import signal
import struct
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from conftest import FIXTURE_GSD, ROOT

import pyglasstools_amd as pg
from pyglasstools_amd import gsd_reader
from pyglasstools_amd.irvingkirkwood import cartesian_grid, polar_grid


# ---------------------------------------------------------------- GSD reader
def test_fixture_header_and_frames():
    t = gsd_reader.open(FIXTURE_GSD)
    assert t.file.application.startswith("gsd.hoomd") and t.file.schema == "hoomd"
    assert t.file.gsd_version == (2, 0)
    assert len(t) == 201
    f = t[180]
    assert f.particles.N == 1024
    assert f.particles.position.shape == (1024, 3) and f.particles.position.dtype == np.float32
    assert f.particles.diameter.dtype == np.float32
    np.testing.assert_allclose(f.configuration.box[:3], [31.84119, 31.84119, 1.0], rtol=1e-7)
    assert f.configuration.step == 180 and t[0].configuration.step == 0
    assert np.all(f.particles.position[:, 2] == 0)
    assert np.all(np.abs(f.particles.position[:, :2]) <= 31.84119 / 2 + 1e-4)
    # chunks the file does not hold take gsd.hoomd's defaults
    assert np.all(f.particles.mass == 1.0) and not f.particles.velocity.any()
    assert 0.72 < f.particles.diameter.min() < f.particles.diameter.max() < 1.62
    t.close()


def _write_gsd(path, frames):
    """Minimal GSD 2.0 writer: frames is a list of {name: ndarray}."""
    names = sorted({k for fr in frames for k in fr})
    ids = {n: i for i, n in enumerate(names)}
    codes = {np.dtype(np.uint8): 1, np.dtype(np.uint32): 3, np.dtype(np.uint64): 4, np.dtype(np.float32): 9,
             np.dtype(np.float64): 10, np.dtype(np.int32): 7}
    nl_n = max(1, (len(names) * 64 + 63) // 64)
    data, entries = bytearray(), []
    off = 256 + nl_n * 64
    for k, fr in enumerate(frames):
        for n, a in fr.items():
            a = np.ascontiguousarray(a)
            N, M = (a.shape[0], a.shape[1] if a.ndim > 1 else 1)
            entries.append((k, N, off + len(data), M, ids[n], codes[a.dtype], 0))
            data += a.tobytes()
    ix_loc = off + len(data)
    hdr = struct.pack("<QQQQQII", 0x65DF65DF65DF65DF, ix_loc, len(entries), 256, nl_n, 0x10004, 0x20000)
    hdr += b"test-writer".ljust(64, b"\0") + b"hoomd".ljust(64, b"\0") + b"\0" * 80
    nl = b"".join(n.encode() + b"\0" for n in names).ljust(nl_n * 64, b"\0")
    with open(path, "wb") as fh:
        fh.write(hdr + nl + bytes(data))
        for e in entries:
            fh.write(struct.pack("<QQqIHBB", *e))


def test_reader_frame0_fallback_and_defaults(tmp_path):
    p = str(tmp_path / "t.gsd")
    pos0 = np.arange(12, dtype=np.float32).reshape(4, 3)
    frames = [
        {"configuration/box": np.array([5, 6, 7, 0, 0, 0], np.float32), "particles/N": np.array([4], np.uint32),
         "particles/position": pos0, "particles/mass": np.full(4, 2.0, np.float32)},
        {"configuration/step": np.array([10], np.uint64), "particles/position": pos0 + 1},
    ]
    _write_gsd(p, frames)
    t = gsd_reader.open(p)
    assert len(t) == 2
    f1 = t[1]
    assert f1.configuration.step == 10 and t[0].configuration.step == 0
    np.testing.assert_array_equal(f1.configuration.box[:3], [5, 6, 7])  # from frame 0
    np.testing.assert_array_equal(f1.particles.mass, 2.0)  # from frame 0
    np.testing.assert_array_equal(f1.particles.position, pos0 + 1)
    np.testing.assert_array_equal(f1.particles.diameter, 1.0)  # schema default
    assert f1.particles.N == 4 and t[-1].configuration.step == 10
    with pytest.raises(IndexError):
        t[2]


def test_reader_rejects_non_gsd(tmp_path):
    p = tmp_path / "bad.gsd"
    p.write_bytes(b"\0" * 300)
    with pytest.raises(gsd_reader.GSDFormatError):
        gsd_reader.open(str(p))


# ---------------------------------------------------------------- grids
def test_cartesian_grid_matches_reference_rule():
    L, dx = 31.84119, 0.25
    g = cartesian_grid(L, dx)
    nmax = int(L / dx)
    pts = np.linspace(-L / 2, L / 2, nmax)
    assert g.shape == (nmax * nmax, 3) == (16129, 3)
    k = 0
    for i in range(nmax):  # x outer, y inner, z = 0 (irvingkirkwood/__init__.py:46-51)
        for j in range(0, nmax, 37):
            assert tuple(g[i * nmax + j]) == (pts[i], pts[j], 0.0)
            k += 1
    assert g[0, 0] == -L / 2 and g[-1, 0] == L / 2


def test_polar_grid_structure():
    gid, r, th, gp = polar_grid(np.zeros(3), rmax=2.0, dr=0.5)
    assert gid[0] == 0 and r[0] == 0 and tuple(gp[0]) == (0, 0, 0)
    rings = np.linspace(0, 2.0, int(2.0 / 0.5))[1:]
    for k, rr in enumerate(rings):
        sel = gid == k + 1
        assert sel.sum() == int(2 * np.pi / (0.5 / rr))
        np.testing.assert_allclose(np.hypot(gp[sel, 0], gp[sel, 1]), rr, rtol=1e-12)


# ---------------------------------------------------------------- pybind surface
def test_simbox_2d_semantics():
    b = pg._pyglasstools.SimBox(np.array([4.0, 5.0, 9.0]), np.zeros(3), 2)
    np.testing.assert_array_equal(b.boxsize, [4, 5, 1])
    np.testing.assert_array_equal(b.periodic, [1, 1, 0])
    np.testing.assert_array_equal(b.upperbound, [2, 2.5, 4.5])  # bounds set before z is reset
    assert b.vol == 20.0
    np.testing.assert_allclose(b.minImage(np.array([3.0, -3.0, 7.0])), [-1.0, 2.0, 7.0])
    b3 = pg._pyglasstools.SimBox(3.0, np.zeros(3), 3)
    assert b3.vol == 27.0


def test_particle_system_validates_sizes():
    b = pg._pyglasstools.SimBox(np.array([4.0, 4.0, 1.0]), np.zeros(3), 2)
    ps = pg._pyglasstools.ParticleSystem(b, 3, np.ones(3), np.ones(3), np.zeros((3, 3)), np.zeros((3, 3)))
    with pytest.raises(ValueError):
        ps.setMass(np.ones(2))
    with pytest.raises(ValueError):
        ps.setAtomPosition(np.zeros((4, 3)))
    with pytest.raises(ValueError):
        pg._pyglasstools.ParticleSystem(b, 3, np.ones(2), np.ones(3), np.zeros((3, 3)), np.zeros((3, 3)))
    ps.setAtomPosition(np.array([[0.1, 0, 0], [1.9, 0, 0], [-1.9, 0, 0]], np.float32))
    assert ps.getNeighbors(np.array([2.0, 0, 0]), 0.5) == [1, 2]  # periodic image
    calc = pg._pyglasstools.Calculator(ps, pg.potential.polydisperse12().cpppairpotential)
    assert calc.max_rcut == pytest.approx(1.25)


def test_potential_coefficients_and_params():
    p = pg.potential.polydisperse12(v0=1.0, eps=0.2, rcut=1.25)
    v0, eps, c0, c1, c2 = p.get_params()
    assert (v0, eps) == (1.0, 0.2)
    assert c0 == -28.0 / 1.25 ** 12 and c1 == 48.0 / 1.25 ** 14 and c2 == -21.0 / 1.25 ** 16
    p.set_params(0.3, "eps")
    assert p.get_params("eps") == 0.3 and p.get_scaledrcut() == 1.25
    assert pg.get_potential() is p
    with pytest.raises(NotImplementedError):
        pg.potential.lj(1.0, 1.0, 2.5, mode="truncated")


def test_cgfunc_construction_rules():
    c = pg.irvingkirkwood.cgfunc.octic(rcut=2.0, mode="fixedpoint", order=5)
    assert c.get_rcut() == 2.0
    c.set_rcut(1.5)
    assert c.get_rcut() == 1.5
    pg.irvingkirkwood.cgfunc.mollifier(rcut=2.0, mode="adaptive", maxiter=100, relerr=1e-8, abserr=1e-10)
    with pytest.raises(TypeError):
        pg.irvingkirkwood.cgfunc.rect(rcut=2.0, mode="fixedpoint")


def test_communicator_is_single_rank():
    c = pg.comm
    assert (c.getRank(), c.getSizeGlobal(), pg.rank, pg.size) == (0, 1, 0, 1)
    a = np.arange(6.0).reshape(2, 3)
    np.testing.assert_array_equal(c.scatter_v([a], 0), a)
    assert c.bcast(7, 0) == 7
    c.barrier()
    c.abort(1)  # no-op on one rank


def test_field_format_is_cpp_stream_default():
    from pyglasstools_amd.irvingkirkwood import _irvingkirkwood

    f = _irvingkirkwood.VirialStressField2D("virialstress", "2-TENSOR", False, 6)
    vals = np.array([1234567.0, 1.2345e-5, -0.0, 0.0, 3.0, 2.0 / 3.0])
    f.values[:, 0, 0] = vals
    for g, v in enumerate(vals):
        assert f.gettostring(0, g) == ("%g " % v)
    assert f.formatColumn(0, 6) == ["%g " % v for v in vals]
    f.clear(0)
    assert f.values[0, 0, 0] == 0.0
    with pytest.raises(IndexError):
        f.gettostring(4, 0)


# ---------------------------------------------------------------- frame loop
DRIVER = textwrap.dedent("""
    import os, sys, signal
    import numpy as np
    sys.path.insert(0, {root!r})
    import pyglasstools as pglass
    import pyglasstools.io as io
    import pyglasstools.utils as utils
    import pyglasstools.potential as pot
    import pyglasstools.irvingkirkwood as ik
    import pyglasstools.irvingkirkwood.cgfunc as cgfunc
    poly12 = pot.polydisperse12(v0=1.0, eps=0.2, rcut=1.25)
    mycgfunc = cgfunc.octic(rcut=2.0, mode="fixedpoint", order=5)
    mysystem = utils.read_gsd(filename={gsd!r}, checkpointfile="restart.log")
    calc = ik.ikcalculator(cgfunc=mycgfunc, dx={dx})
    names = ['virialstress_00', 'virialstress_11', 'virialstress_01']
    logger = io.fieldlogger(keyword="test-dump", names=names, solver=calc)
    # stand-in for the GPU compute: a value that identifies the frame
    def fake_run(frame=[None]):
        f = pglass.get_sysdata().frame_num
        v = logger.field_obs["virialstress"].values
        v[:, 0, 0] = f; v[:, 1, 1] = -f; v[:, 0, 1] = v[:, 1, 0] = 0.5
        if f == {stop_at}:
            os.kill(os.getpid(), signal.SIGINT)
    calc.run = fake_run
    pglass.analyze(frame_list=range({f0}, {f1}))
""")


def _run_driver(tmp_path, f0, f1, stop_at=-1, dx=4.0):
    script = tmp_path / "drive.py"
    script.write_text(DRIVER.format(root=ROOT, gsd=FIXTURE_GSD, dx=dx, f0=f0, f1=f1, stop_at=stop_at))
    return subprocess.run([sys.executable, str(script)], cwd=str(tmp_path), capture_output=True, text=True,
                          timeout=120)


def _frames_in(path):
    txt = open(path).read().splitlines()
    frames, k = [], 0
    while k < len(txt):
        G = int(txt[k].split()[0])
        assert txt[k + 1].startswith("#Frame ")
        frames.append((int(txt[k + 1].split()[1]), txt[k + 2:k + 2 + G]))
        k += 2 + G
    return frames


def test_xyz_output_and_checkpoint(tmp_path):
    r = _run_driver(tmp_path, 180, 183)
    assert r.returncode == 0, r.stderr
    frames = _frames_in(tmp_path / "test-dump.xyz")
    assert [f for f, _ in frames] == [180, 181, 182]
    L = 31.84119
    n = int(L / 4.0)
    for f, rows in frames:
        assert len(rows) == n * n
        first = rows[0].split(" ")
        assert first[0] == "1" and float(first[1]) == pytest.approx(-L / 2, rel=1e-6)
        assert rows[0].endswith("%g %g %g " % (f, -f, 0.5))
    assert open(tmp_path / "restart.log").read() == "Frame 183"


def test_resume_from_checkpoint(tmp_path):
    (tmp_path / "restart.log").write_text("Frame 181")
    r = _run_driver(tmp_path, 180, 184)
    assert r.returncode == 0, r.stderr
    assert [f for f, _ in _frames_in(tmp_path / "test-dump.xyz")] == [181, 182, 183]
    assert open(tmp_path / "restart.log").read() == "Frame 184"
    # nothing left to do
    r = _run_driver(tmp_path, 180, 184)
    assert "No more frames" in r.stdout


def test_interrupt_completes_the_frame_then_stops(tmp_path):
    # SIGINT raised inside frame 181's compute: that frame is still logged and
    # checkpointed (the signal is held during the save), then the run stops
    r = _run_driver(tmp_path, 180, 185, stop_at=181)
    assert r.returncode != 0 and "KeyboardInterrupt" in r.stderr
    frames = [f for f, _ in _frames_in(tmp_path / "test-dump.xyz")]
    assert frames in ([180], [180, 181])
    assert open(tmp_path / "restart.log").read() == "Frame %d" % (frames[-1] + 1)
    # and a rerun resumes exactly where the log says
    r = _run_driver(tmp_path, 180, 185)
    assert r.returncode == 0, r.stderr
    assert [f for f, _ in _frames_in(tmp_path / "test-dump.xyz")] == list(range(180, 185))


def test_autotuner_picks_fastest_shape_and_caches(tmp_path):
    """The autotuner times every (gather kind, cell reach, points per lane,
    threads, frames per launch) shape on the caller's job -- the tile gather
    only over reach and frames per launch, its workgroup shape being fixed --
    keeps the fastest, and replays it from its cache (in memory and on disk)
    without timing again."""
    from pyglasstools_amd.autotune import Autotuner

    class FakeEngine(object):
        def __init__(self):
            self.calls = []
            self.shape = None
            self.mode = "lane"

        def set_tuning(self, threads_per_block=0, frames_per_launch=0, points_per_lane=0, cell_reach=0):
            self.calls.append((self.mode, threads_per_block, frames_per_launch, points_per_lane, cell_reach))
            self.shape = (threads_per_block, frames_per_launch, points_per_lane, cell_reach)

        def set_gather_mode(self, mode):
            self.mode = mode

    def run(cost, fast):
        eng = FakeEngine()

        def job():
            import time
            time.sleep(cost.get((eng.mode,) + eng.shape, 2e-3))

        path = str(tmp_path / ("tune_%s.json" % fast[0]))
        best, trials = Autotuner(path).tune(eng, job, ("k", 1), 40, fpl=(10, 100), repeats=1)
        assert (best["gather"], best["threads_per_block"], best["frames_per_launch"], best["points_per_lane"],
                best["cell_reach"]) == fast
        # lane: 2 reaches x 3 points-per-lane x 2 blocks x 2 frame counts (100 is
        # capped at the 40 frames); tile: 2 reaches x 2 frame counts
        assert len(trials) == 28 and {t["frames_per_launch"] for t in trials} == {10, 40}
        assert sum(t["gather"] == "tile" for t in trials) == 4
        assert (eng.mode,) + eng.shape == fast
        eng2 = FakeEngine()
        best2, trials2 = Autotuner(path).tune(eng2, job, ("k", 1), 40)
        assert best2 == best and trials2 == [] and eng2.calls == [("lane",) + fast[1:]] and eng2.mode == fast[0]

    run({("lane", 128, 10, 2, 2): 1e-4}, ("lane", 128, 10, 2, 2))
    run({("tile", 128, 40, 1, 1): 1e-4}, ("tile", 128, 40, 1, 1))
