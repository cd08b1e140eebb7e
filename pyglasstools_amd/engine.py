"""ctypes binding of the C ABI in include/pgt_ik.h (libpgt_ik.so).

This is the boundary the tests and the benchmark call: plain pointers and
sizes, host numpy arrays for pgt_ik_compute, device pointers (any HIP
allocation, e.g. torch tensors' data_ptr()) for pgt_ik_compute_frames.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpgt_ik.so")

PGT_OK = 0
ERRORS = {-1: "PGT_ERR_ARG", -2: "PGT_ERR_HIP", -3: "PGT_ERR_STATE", -4: "PGT_ERR_RANGE", -5: "PGT_ERR_UNSUPPORTED"}
POT = {"poly12": 0, "poly18": 1, "poly10": 2, "poly106": 3, "polylj": 4}
CG = {"octic": 0, "rect": 1, "mollifier": 2}
FIELD_VIRIAL, FIELD_KINETIC, FIELD_ELASTIC = 1, 2, 4
PREC = {"fp64": 0, "fp32": 1}
MLP_IN_PAD = 8  # floats per network input row (kMlpInPad)

# every symbol include/pgt_ik.h declares
EXPORTED = [
    "pgt_last_error", "pgt_device_count", "pgt_ik_create", "pgt_ik_destroy", "pgt_ik_set_box",
    "pgt_ik_set_particles", "pgt_ik_set_grid", "pgt_ik_compute", "pgt_ik_compute_frames",
    "pgt_ik_compute_frames_noalloc", "pgt_ik_enqueue_frames",
    "pgt_ik_finish", "pgt_ik_release_enqueued", "pgt_ik_set_stream",
    "pgt_ik_set_tuning", "pgt_ik_set_gather_mode", "pgt_ik_get_stats", "pgt_eval_pair_force", "pgt_eval_cgfunc",
    "pgt_mlp_nparam", "pgt_mlp_work_floats", "pgt_mlp_forward", "pgt_mlp_grad", "pgt_mlp_adam", "pgt_mlp_infer",
    "pgt_ik_bond_samples", "pgt_ik_compute_surrogate", "pgt_pretest_band",
    "pgt_ik_sweep_create", "pgt_ik_sweep_destroy", "pgt_ik_sweep_run", "pgt_ik_sweep_enqueue",
    "pgt_ik_sweep_release", "pgt_ik_sweep_finish", "pgt_ik_sweep_set_stream", "pgt_ik_sweep_get_stats",
]


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("potential_kind", ctypes.c_int),
        ("scaled_rcut", ctypes.c_double),
        ("pot_params", ctypes.c_double * 6),
        ("n_pot_params", ctypes.c_int),
        ("cg_kind", ctypes.c_int),
        ("cg_rcut", ctypes.c_double),
        ("quad_mode", ctypes.c_int),
        ("quad_order", ctypes.c_int),
        ("adapt_limit", ctypes.c_int),
        ("adapt_epsrel", ctypes.c_double),
        ("adapt_epsabs", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("max_rcut", ctypes.c_double),
    ]


class SweepSystem(ctypes.Structure):
    _fields_ = [
        ("engine", ctypes.c_void_p),
        ("nframes", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("d_pos", ctypes.c_void_p),
        ("d_diam", ctypes.c_void_p),
        ("d_mass", ctypes.c_void_p),
        ("d_vel", ctypes.c_void_p),
        ("d_out_virial", ctypes.c_void_p),
        ("d_out_kinetic", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_particles", ctypes.c_int64),
        ("n_grid", ctypes.c_int64),
        ("n_ball", ctypes.c_int64),
        ("n_contrib", ctypes.c_int64),
        ("bond_capacity", ctypes.c_int32),
        ("max_bonds", ctypes.c_int32),
        ("gather_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("n_pairs", ctypes.c_int64),
        ("n_network", ctypes.c_int64),
        ("surrogate_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("%s is not built; run `python -m pyglasstools_amd.csrc.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        L.pgt_last_error.restype = ctypes.c_char_p
        L.pgt_device_count.restype = i
        L.pgt_pretest_band.argtypes = [d, d, d, d, i, i, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float)]
        L.pgt_ik_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.pgt_ik_destroy.argtypes = [vp]
        L.pgt_ik_destroy.restype = None
        L.pgt_ik_set_box.argtypes = [vp, vp, vp, vp]
        L.pgt_ik_set_particles.argtypes = [vp, i64, vp, vp, vp, vp, vp]
        L.pgt_ik_set_grid.argtypes = [vp, i64, vp]
        L.pgt_ik_compute.argtypes = [vp, ctypes.c_uint, vp, vp, vp]
        L.pgt_ik_compute_frames.argtypes = [vp, i, i64, vp, vp, vp, vp, vp, ctypes.c_uint, vp, vp, vp]
        L.pgt_ik_enqueue_frames.argtypes = [vp, i, i64, vp, vp, vp, vp, vp, ctypes.c_uint, vp, vp, vp]
        L.pgt_ik_compute_frames_noalloc.argtypes = [vp, i, i64, vp, vp, vp, vp, vp, ctypes.c_uint, vp, vp, vp]
        L.pgt_ik_finish.argtypes = [vp]
        L.pgt_ik_release_enqueued.argtypes = [vp]
        L.pgt_ik_set_stream.argtypes = [vp, vp]
        L.pgt_ik_set_tuning.argtypes = [vp, i, i, i, i]
        L.pgt_ik_set_gather_mode.argtypes = [vp, i]
        L.pgt_ik_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.pgt_eval_pair_force.argtypes = [i, vp, i, d, i64, vp, vp, vp, vp, vp, vp]
        L.pgt_eval_cgfunc.argtypes = [i, d, i, i, i, d, d, i, i64, vp, vp, vp, vp, vp]
        L.pgt_mlp_nparam.argtypes = []
        L.pgt_mlp_work_floats.argtypes = [i64]
        L.pgt_mlp_forward.argtypes = [vp, i64, vp, vp, vp, vp]
        L.pgt_mlp_grad.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp]
        L.pgt_mlp_adam.argtypes = [vp, vp, vp, vp, i, ctypes.c_float, vp]
        L.pgt_mlp_infer.argtypes = [vp, i64, vp, vp, vp]
        L.pgt_ik_bond_samples.argtypes = [vp, i64, vp, vp, ctypes.POINTER(i64)]
        L.pgt_ik_compute_surrogate.argtypes = [vp, vp, d, d, d, vp]
        L.pgt_ik_sweep_create.argtypes = [i, ctypes.POINTER(SweepSystem), ctypes.c_uint, vp, ctypes.POINTER(vp)]
        L.pgt_ik_sweep_destroy.argtypes = [vp]
        for name in ("pgt_ik_sweep_run", "pgt_ik_sweep_enqueue", "pgt_ik_sweep_release", "pgt_ik_sweep_finish"):
            getattr(L, name).argtypes = [vp]
        L.pgt_ik_sweep_set_stream.argtypes = [vp, vp]
        L.pgt_ik_sweep_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        for name in EXPORTED:
            if name not in ("pgt_last_error", "pgt_device_count", "pgt_ik_destroy", "pgt_ik_sweep_destroy"):
                getattr(L, name).restype = i
        L.pgt_mlp_work_floats.restype = ctypes.c_size_t
        _LIB = L
    return _LIB


class PGTError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = lib().pgt_last_error().decode(errors="replace")
        super().__init__("%s: %s (%s)" % (where, msg, ERRORS.get(code, code)))


def _check(rc, where):
    if rc != PGT_OK:
        raise PGTError(rc, where)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    if hasattr(a, "data_ptr"):  # torch tensor on the device
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _dev(a, dtype, numel, name):
    """Device pointer of a torch tensor after checking what the kernels
    assume: on the GPU, contiguous, of `dtype`, holding at least `numel`
    elements.  None passes through.  Anything else raises ValueError before
    any launch: the C side cannot check sizes, and a host, strided, narrower
    or shorter buffer would be read or written out of bounds on the device."""
    if a is None:
        return None
    import torch

    if not isinstance(a, torch.Tensor):
        raise ValueError("%s must be a torch tensor on the GPU (got %s)" % (name, type(a).__name__))
    if not a.is_cuda:
        raise ValueError("%s must be on the GPU (got device %s)" % (name, a.device))
    if a.dtype != dtype:
        raise ValueError("%s must be %s (got %s)" % (name, dtype, a.dtype))
    if not a.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if a.numel() < numel:
        raise ValueError("%s holds %d elements; this call reads or writes %d" % (name, a.numel(), numel))
    return ctypes.c_void_p(a.data_ptr())


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def device_count():
    return lib().pgt_device_count()


def pretest_band(r, max_coord, max_box, min_box, dim=3, periodic=True):
    """(in32, out32): the FP32 pre-test thresholds the engine sets for search
    radius r (pgt_pretest_band); host code, no device needed."""
    lo, hi = ctypes.c_float(), ctypes.c_float()
    _check(lib().pgt_pretest_band(float(r), float(max_coord), float(max_box), float(min_box), int(dim),
                                  1 if periodic else 0, ctypes.byref(lo), ctypes.byref(hi)), "pgt_pretest_band")
    return np.float32(lo.value), np.float32(hi.value)


def poly12_params(v0=1.0, eps=0.2, rcut=1.25):
    return [v0, eps, -28.0 * v0 / rcut ** 12, 48.0 * v0 / rcut ** 14, -21.0 * v0 / rcut ** 16]


class Engine(object):
    """One pgt_ik_engine."""

    def __init__(self, dim=2, pot="poly12", params=None, scaled_rcut=1.25, cg="octic", cg_rcut=2.0, order=5,
                 max_rcut=None, precision="fp64", diameters=None, adaptive=None):
        """adaptive: None for the fixed-order Gauss-Legendre rule of `order`
        nodes, else (limit, epsrel, epsabs) of the QAG rule (AdaptiveCGFunc)."""
        if params is None:
            params = poly12_params(rcut=scaled_rcut)
        if max_rcut is None:
            if diameters is None:
                raise ValueError("give max_rcut or the diameters it is computed from")
            dmax = float(np.max(diameters))
            max_rcut = max(scaled_rcut * dmax, dmax)
        p = Params()
        p.dim = dim
        p.potential_kind = POT[pot]
        p.scaled_rcut = scaled_rcut
        p.n_pot_params = len(params)
        for k, v in enumerate(params):
            p.pot_params[k] = v
        p.cg_kind = CG[cg]
        p.cg_rcut = cg_rcut
        if adaptive is None:
            p.quad_mode = 0
            p.quad_order = order
        else:
            p.quad_mode = 1
            p.quad_order = 0
            p.adapt_limit, p.adapt_epsrel, p.adapt_epsabs = int(adaptive[0]), float(adaptive[1]), float(adaptive[2])
        p.precision = PREC[precision]
        p.max_rcut = max_rcut
        self.params = p
        self.dim = dim
        self.max_rcut = max_rcut
        h = ctypes.c_void_p()
        _check(lib().pgt_ik_create(ctypes.byref(p), ctypes.byref(h)), "pgt_ik_create")
        self.h = h
        self.G = 0

    def close(self):
        if self.h:
            lib().pgt_ik_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_box(self, L, periodic, lower=None):
        L = _f64(L)
        lo = _f64(lower) if lower is not None else -0.5 * L
        per = np.ascontiguousarray(periodic, dtype=np.int32)
        _check(lib().pgt_ik_set_box(self.h, _ptr(L), _ptr(lo), _ptr(per)), "pgt_ik_set_box")

    def set_particles(self, pos, diam, mass=None, vel=None, disp=None):
        pos = _f64(pos, (-1, 3))
        n = pos.shape[0]
        self._keep = [pos, _f64(diam), None if mass is None else _f64(mass), None if vel is None else _f64(vel, (-1, 3)),
                      None if disp is None else _f64(disp, (-1, 3))]
        _check(lib().pgt_ik_set_particles(self.h, n, *[_ptr(a) for a in self._keep]), "pgt_ik_set_particles")

    def set_grid(self, grid):
        grid = _f64(grid, (-1, 3))
        self.G = grid.shape[0]
        _check(lib().pgt_ik_set_grid(self.h, self.G, _ptr(grid)), "pgt_ik_set_grid")

    def set_system(self, system, grid):
        self.set_box(system["L"], system["periodic"])
        self.set_particles(system["pos"], system["diam"], system.get("mass"), system.get("vel"), system.get("disp"))
        self.set_grid(grid)

    def bond_samples(self, d_x=None, d_t=None, max_pairs=0):
        """Training samples of the current frame and grid (device float32
        tensors: d_x rows of 8, d_t the 5-node integrals); returns the pair
        count (only counts when d_x is None)."""
        import torch

        n = ctypes.c_int64(0)
        max_pairs = int(max_pairs)
        px = _dev(d_x, torch.float32, MLP_IN_PAD * max_pairs, "d_x")
        pt = _dev(d_t, torch.float32, max_pairs, "d_t")
        _check(lib().pgt_ik_bond_samples(self.h, max_pairs, px, pt, ctypes.byref(n)),
               "pgt_ik_bond_samples")
        return n.value

    def compute_surrogate(self, d_params, bound_inside, bound_crossing, tol):
        """Virial field (G, 2, 2) with the network's bond integral where the
        pair class's error bound is within tol, the 5-node rule elsewhere."""
        import torch

        out = np.zeros((self.G, 2, 2))
        pp = _dev(d_params, torch.float32, mlp_nparam(), "d_params")
        _check(lib().pgt_ik_compute_surrogate(self.h, pp, bound_inside, bound_crossing, tol, _ptr(out)),
               "pgt_ik_compute_surrogate")
        return out

    def compute(self, fields=("virial",)):
        D = self.dim
        out = {}
        mask = 0
        ptrs = [None, None, None]
        for k, (name, bit) in enumerate((("virial", FIELD_VIRIAL), ("kinetic", FIELD_KINETIC), ("elastic", FIELD_ELASTIC))):
            if name in fields:
                out[name] = np.zeros((self.G, D, D))
                ptrs[k] = _ptr(out[name])
                mask |= bit
        _check(lib().pgt_ik_compute(self.h, mask, *ptrs), "pgt_ik_compute")
        return out

    def _frame_args(self, nframes, n, d_pos, d_diam, d_mass, d_vel, d_disp, fields, d_virial, d_kinetic, d_elastic):
        """Field mask and checked device pointers of a batched call."""
        import torch

        nframes, n = int(nframes), int(n)
        if nframes < 0 or n < 0:
            raise ValueError("frame and particle counts must be non-negative")
        mask = (FIELD_VIRIAL if "virial" in fields else 0) | (FIELD_KINETIC if "kinetic" in fields else 0) | \
               (FIELD_ELASTIC if "elastic" in fields else 0)
        f64 = torch.float64
        m = nframes * n
        fsz = nframes * self.G * self.dim * self.dim
        if d_pos is None or d_diam is None:
            raise ValueError("positions and diameters are required")
        ptrs = [_dev(d_pos, f64, 3 * m, "d_pos"), _dev(d_diam, f64, m, "d_diam"), _dev(d_mass, f64, m, "d_mass"),
                _dev(d_vel, f64, 3 * m, "d_vel"), _dev(d_disp, f64, 3 * m, "d_disp")]
        outs = []
        for name, bit, a in (("d_virial", FIELD_VIRIAL, d_virial), ("d_kinetic", FIELD_KINETIC, d_kinetic),
                             ("d_elastic", FIELD_ELASTIC, d_elastic)):
            if mask & bit and a is None:
                raise ValueError("%s is required for the selected fields" % name)
            outs.append(_dev(a, f64, fsz, name) if mask & bit else None)
        return nframes, n, mask, ptrs, outs

    def compute_frames(self, nframes, n, d_pos, d_diam, d_mass=None, d_vel=None, d_disp=None, fields=("virial",),
                       d_virial=None, d_kinetic=None, d_elastic=None):
        """Fields of nframes device-resident frames (torch float64 tensors
        on the GPU, frames back to back); outputs nframes x G x D x D."""
        nframes, n, mask, ptrs, outs = self._frame_args(nframes, n, d_pos, d_diam, d_mass, d_vel, d_disp, fields,
                                                        d_virial, d_kinetic, d_elastic)
        _check(lib().pgt_ik_compute_frames(self.h, nframes, n, *ptrs, mask, *outs), "pgt_ik_compute_frames")

    def compute_frames_noalloc(self, nframes, n, d_pos, d_diam, d_mass=None, d_vel=None, d_disp=None,
                               fields=("virial",), d_virial=None, d_kinetic=None, d_elastic=None):
        """compute_frames with every work array in the library's static
        device scratch array: the call allocates nothing."""
        nframes, n, mask, ptrs, outs = self._frame_args(nframes, n, d_pos, d_diam, d_mass, d_vel, d_disp, fields,
                                                        d_virial, d_kinetic, d_elastic)
        _check(lib().pgt_ik_compute_frames_noalloc(self.h, nframes, n, *ptrs, mask, *outs),
               "pgt_ik_compute_frames_noalloc")

    def set_stream(self, stream_handle):
        _check(lib().pgt_ik_set_stream(self.h, ctypes.c_void_p(stream_handle) if stream_handle else None),
               "pgt_ik_set_stream")

    def enqueue_frames(self, nframes, n, d_pos, d_diam, d_mass=None, d_vel=None, d_disp=None, fields=("virial",),
                       d_virial=None, d_kinetic=None, d_elastic=None):
        """compute_frames without host synchronisation or allocation (graph
        capture); needs an earlier compute_frames of the same shape; check
        with finish()."""
        nframes, n, mask, ptrs, outs = self._frame_args(nframes, n, d_pos, d_diam, d_mass, d_vel, d_disp, fields,
                                                        d_virial, d_kinetic, d_elastic)
        _check(lib().pgt_ik_enqueue_frames(self.h, nframes, n, *ptrs, mask, *outs), "pgt_ik_enqueue_frames")

    def finish(self):
        _check(lib().pgt_ik_finish(self.h), "pgt_ik_finish")

    def release_enqueued(self):
        """No captured graph replays this engine's enqueued launches any
        more: calls may regrow its buffers again."""
        _check(lib().pgt_ik_release_enqueued(self.h), "pgt_ik_release_enqueued")

    def set_tuning(self, threads_per_block=0, frames_per_launch=0, points_per_lane=0, cell_reach=0):
        _check(lib().pgt_ik_set_tuning(self.h, threads_per_block, frames_per_launch, points_per_lane, cell_reach),
               "pgt_ik_set_tuning")

    def set_gather_mode(self, mode):
        """'tile' (one wavefront per 64-point tile, work items balanced by
        bond count) or 'lane' (one lane per grid point or group of points)
        for the Octic 5-node rule; other rules always run the lane gather."""
        _check(lib().pgt_ik_set_gather_mode(self.h, {"lane": 0, "tile": 1}[mode]), "pgt_ik_set_gather_mode")

    def stats(self):
        s = Stats()
        _check(lib().pgt_ik_get_stats(self.h, ctypes.byref(s)), "pgt_ik_get_stats")
        return s.as_dict()


class Sweep(object):
    """One pgt_ik_sweep: the frames of several systems (all 2D or all 3D),
    one launch per stage over all of them.  systems: dicts with the system's
    Engine (box, grid and parameters set) under "engine", "nframes", "n", the
    device inputs "pos", "diam" (and optionally "mass", "vel"; torch float64
    tensors, frames back to back) and the outputs "virial" / "kinetic"
    (nframes x G x D x D) for the selected fields.  The sweep keeps the
    tensors alive and reads and writes them on every run."""

    def __init__(self, systems, fields=("virial", "kinetic"), stream_handle=None):
        import torch

        mask = (FIELD_VIRIAL if "virial" in fields else 0) | (FIELD_KINETIC if "kinetic" in fields else 0) | \
               (FIELD_ELASTIC if "elastic" in fields else 0)
        f64 = torch.float64
        arr = (SweepSystem * len(systems))()
        self._keep = []
        for k, y in enumerate(systems):
            e, nf, n = y["engine"], int(y["nframes"]), int(y["n"])
            if nf < 1 or n < 1:
                raise ValueError("system %d: frame and particle counts must be positive" % k)
            m, fsz = nf * n, nf * e.G * e.dim * e.dim
            a = arr[k]
            a.engine = e.h.value if isinstance(e.h, ctypes.c_void_p) else e.h
            a.nframes, a.n = nf, n
            a.d_pos = _dev(y["pos"], f64, 3 * m, "pos").value
            a.d_diam = _dev(y["diam"], f64, m, "diam").value
            p = _dev(y.get("mass"), f64, m, "mass")
            a.d_mass = p.value if p is not None else None
            p = _dev(y.get("vel"), f64, 3 * m, "vel")
            a.d_vel = p.value if p is not None else None
            for name, bit in (("virial", FIELD_VIRIAL), ("kinetic", FIELD_KINETIC)):
                if mask & bit:
                    if y.get(name) is None:
                        raise ValueError("system %d: %s output is required for the selected fields" % (k, name))
                    setattr(a, "d_out_" + name, _dev(y[name], f64, fsz, name).value)
            self._keep.append(y)
        self.systems = list(systems)
        h = ctypes.c_void_p()
        _check(lib().pgt_ik_sweep_create(len(systems), arr, mask,
                                         ctypes.c_void_p(stream_handle) if stream_handle else None, ctypes.byref(h)),
               "pgt_ik_sweep_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _LIB is not None:
            _LIB.pgt_ik_sweep_destroy(h)
            self.h = None

    def run(self):
        _check(lib().pgt_ik_sweep_run(self.h), "pgt_ik_sweep_run")

    def enqueue(self):
        """The sweep's launches, no synchronisation or allocation (graph
        capture); needs an earlier run(); check with finish()."""
        _check(lib().pgt_ik_sweep_enqueue(self.h), "pgt_ik_sweep_enqueue")

    def release(self):
        _check(lib().pgt_ik_sweep_release(self.h), "pgt_ik_sweep_release")

    def finish(self):
        _check(lib().pgt_ik_sweep_finish(self.h), "pgt_ik_sweep_finish")

    def set_stream(self, stream_handle):
        _check(lib().pgt_ik_sweep_set_stream(self.h, ctypes.c_void_p(stream_handle) if stream_handle else None),
               "pgt_ik_sweep_set_stream")

    def stats(self):
        s = Stats()
        _check(lib().pgt_ik_sweep_get_stats(self.h, ctypes.byref(s)), "pgt_ik_sweep_get_stats")
        return s.as_dict()


def eval_pair_force(rij, di, dj, pot="poly12", params=None, scaled_rcut=1.25):
    if params is None:
        params = poly12_params(rcut=scaled_rcut)
    rij = _f64(rij, (-1, 3))
    n = rij.shape[0]
    di, dj = _f64(np.broadcast_to(di, (n,))), _f64(np.broadcast_to(dj, (n,)))
    p = _f64(params)
    f, e, r = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(lib().pgt_eval_pair_force(POT[pot], _ptr(p), len(params), scaled_rcut, n, _ptr(rij), _ptr(di), _ptr(dj),
                                     _ptr(f), _ptr(e), _ptr(r)), "pgt_eval_pair_force")
    return f, e, r


def eval_cgfunc(mode, x, ri, rij=None, s=None, cg="octic", rcut=2.0, order=5, adaptive=None):
    """Kernel (delta), integrand (obj) or bond integral (bond) on the GPU;
    adaptive = (limit, epsrel, epsabs) selects the QAG rule for the bond."""
    x = _f64(x, (-1, 3))
    n = x.shape[0]
    ri = _f64(ri, (-1, 3))
    rij = None if rij is None else _f64(rij, (-1, 3))
    s = None if s is None else _f64(np.broadcast_to(s, (n,)))
    out = np.zeros(n)
    modes = {"delta": 0, "obj": 1, "bond": 2}
    qm, lim, er, ea = (0, 0, 0.0, 0.0) if adaptive is None else (1, int(adaptive[0]), float(adaptive[1]),
                                                                 float(adaptive[2]))
    _check(lib().pgt_eval_cgfunc(CG[cg], rcut, qm, order, lim, er, ea, modes[mode], n, _ptr(s), _ptr(x), _ptr(ri),
                                 _ptr(rij), _ptr(out)), "pgt_eval_cgfunc")
    return out


# ---------------------------------------------------------------------------
# bond-integral surrogate network (device tensors; work on torch's current
# stream so it orders with the caller's own device work)
# ---------------------------------------------------------------------------
def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def mlp_nparam():
    return lib().pgt_mlp_nparam()


def mlp_work_floats(n):
    return int(lib().pgt_mlp_work_floats(n))


def _rows(x):
    """Sample count of a network input (rows of MLP_IN_PAD floats)."""
    if getattr(x, "dim", lambda: 0)() != 2 or x.shape[1] != MLP_IN_PAD:
        raise ValueError("network inputs are rows of %d floats (got shape %s)" % (MLP_IN_PAD, tuple(x.shape)))
    return int(x.shape[0])


def _net(params):
    import torch

    return _dev(params, torch.float32, mlp_nparam(), "params")


def mlp_forward(params, x, y, work):
    import torch

    n = _rows(x)
    f = torch.float32
    _check(lib().pgt_mlp_forward(_net(params), n, _dev(x, f, n * MLP_IN_PAD, "x"), _dev(y, f, n, "y"),
                                 _dev(work, f, mlp_work_floats(n), "work"), _stream()), "pgt_mlp_forward")


def mlp_grad(params, x, t, grad, loss, work):
    import torch

    n = _rows(x)
    f = torch.float32
    _check(lib().pgt_mlp_grad(_net(params), n, _dev(x, f, n * MLP_IN_PAD, "x"), _dev(t, f, n, "t"),
                              _dev(grad, f, mlp_nparam(), "grad"), _dev(loss, f, 1, "loss"),
                              _dev(work, f, mlp_work_floats(n), "work"), _stream()), "pgt_mlp_grad")


def mlp_adam(params, m, v, grad, step, lr):
    import torch

    f, k = torch.float32, mlp_nparam()
    _check(lib().pgt_mlp_adam(_net(params), _dev(m, f, k, "m"), _dev(v, f, k, "v"), _dev(grad, f, k, "grad"), step,
                              lr, _stream()), "pgt_mlp_adam")


def mlp_infer(params, x, y):
    import torch

    n = _rows(x)
    f = torch.float32
    _check(lib().pgt_mlp_infer(_net(params), n, _dev(x, f, n * MLP_IN_PAD, "x"), _dev(y, f, n, "y"), _stream()),
           "pgt_mlp_infer")
