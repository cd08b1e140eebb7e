"""Autotuner for the field-gather launch shape.

The knobs of the engine tuned per workload: the gather kind ("lane": one
lane per grid point or group of points; "tile": one wavefront per 64-point
tile with the work items balanced by bond count), threads per workgroup of
the gather kernel (128 or 256; with the per-lane LDS lists this sets how many
workgroups share a CU), frames per launch (how many frames one kernel launch
covers; more frames amortise launch gaps, fewer keep the work arrays small)
grid points per gather lane (1, 2 or 4: more points share each read of a
particle's bond entries but keep more registers and a longer LDS list), and
the cell reach (1 or 2: cells of the search radius searched one to each side,
or of half of it searched two to each side).  The tile gather has a fixed
workgroup shape, so for it only frames per launch and the reach are tried.
Each candidate is timed on the caller's own job; the best is cached
per workload key in memory and, when a path is given, in a JSON file.
"""
import json
import os
import time

DEFAULT_BLOCK = (128, 256)
DEFAULT_FPL = (4, 10, 25, 50, 100)
DEFAULT_PL = (1, 2, 4)
DEFAULT_REACH = (1, 2)
DEFAULT_GATHER = ("lane", "tile")
KNOBS = ("threads_per_block", "frames_per_launch", "points_per_lane", "cell_reach")


def _key(parts):
    return "|".join(str(p) for p in parts)


class Autotuner(object):
    def __init__(self, cache_path=None):
        self.cache_path = cache_path
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    self.cache = json.load(f)
            except (OSError, ValueError):
                self.cache = {}

    def tune(self, engine, run_job, key, nframes, blocks=DEFAULT_BLOCK, fpl=DEFAULT_FPL, sync=None, repeats=2,
             points=DEFAULT_PL, reaches=DEFAULT_REACH, gathers=DEFAULT_GATHER):
        """run_job() runs the whole job once; sync() waits for the device.
        Returns (best settings dict, list of trial records)."""
        k = _key(key)
        if k in self.cache:
            best = self.cache[k]
            engine.set_tuning(best["threads_per_block"], best["frames_per_launch"], best.get("points_per_lane", 1),
                              best.get("cell_reach", 1))
            engine.set_gather_mode(best.get("gather", "lane"))
            return best, []
        trials = []
        for mode in gathers:
            engine.set_gather_mode(mode)
            shapes = [(b, pl) for pl in points for b in blocks] if mode == "lane" else [(blocks[0], 1)]
            for rc in reaches:
                for b, pl in shapes:
                    for fp in sorted(set(min(f, nframes) for f in fpl)):
                        engine.set_tuning(b, fp, pl, rc)
                        run_job()  # first call allocates for this shape
                        if sync:
                            sync()
                        t = []
                        for _ in range(repeats):
                            t0 = time.perf_counter()
                            run_job()
                            if sync:
                                sync()
                            t.append(time.perf_counter() - t0)
                        trials.append({"gather": mode, "threads_per_block": b, "frames_per_launch": fp,
                                       "points_per_lane": pl, "cell_reach": rc, "seconds": min(t)})
        best = min(trials, key=lambda r: r["seconds"])
        best = {k: best[k] for k in KNOBS + ("gather",)}
        engine.set_tuning(*[best[k] for k in KNOBS])
        engine.set_gather_mode(best["gather"])
        self.cache[k] = best
        if self.cache_path:
            try:
                d = os.path.dirname(self.cache_path)
                if d:
                    os.makedirs(d, exist_ok=True)
                with open(self.cache_path, "w") as f:
                    json.dump(self.cache, f, indent=1, sort_keys=True)
            except OSError:
                pass
        return best, trials
