"""Field output (mirrors the IK part of pyglasstools/io.py).

`fieldlogger` writes `<keyword>.xyz` in the reference's format (io.py:82-169):
per frame a line "{G} ", a line "#Frame {f}  ", then per grid point
"1 x y [z] v0 v1 ... " where coordinates use Python's float formatting and
values the C++ stream default (6 significant digits, CoarseGrainedField.h:178).
The thermo `logfile` (io.py:9-80) belongs to the thermo calculator, which this
package does not provide.
"""
import numpy as np

from . import _pyglasstools, comm, rank, size, loggers_list, get_sysdata
from . import irvingkirkwood
from . import savemode


class logfile(object):
    def __init__(self, filename, names, solver):
        raise NotImplementedError("io.logfile saves thermo observables; the thermo calculator is not part of "
                                  "this package (it provides the Irving-Kirkwood field path)")


class fieldlogger(object):
    def __init__(self, keyword, names, solver):
        self.__obs_names = names
        if solver is None:
            if rank == 0:
                print("[ERROR] User must specify solver object!")
            comm.abort(1)
            raise ValueError("fieldlogger needs a solver")
        elif names is None:
            if rank == 0:
                print("[ERROR] User must specify list of observables to save!")
            comm.abort(1)
            raise ValueError("fieldlogger needs observable names")
        self.solver = solver
        loggers_list.append(self)
        self.field_obs = irvingkirkwood.initialize_field(self.__obs_names, get_sysdata().pysimbox.dim,
                                                         self.solver.gridpoints)
        self.solver.add_observables(self.field_obs)
        self.file = _pyglasstools.MPILogFile(comm, "{}".format(keyword) + ".xyz")
        self._prefix = None

    def _columns(self, Dim):
        cols = []
        for name in self.__obs_names:
            i = int(name[-2])
            j = int(name[-1])
            cols.append((self.field_obs[name[:-3]], i + Dim * j))
        return cols

    def _prefixes(self, Dim, polar):
        """Per-grid-point line starts; they do not change between frames."""
        if self._prefix is None:
            out = []
            if polar:
                for index in range(len(self.solver.gridpoints)):
                    out.append("{} {} {} {} ".format(1, int(self.solver.gridid[index]), self.solver.globalr[index],
                                                      self.solver.globaltheta[index]))
            else:
                for index in range(len(self.solver.gridpoints)):
                    g = self.solver.gridpoints[index]
                    line = "{} {} {} ".format(1, g[0], g[1])
                    if Dim == 3:
                        line += "{} ".format(g[2])
                    out.append(line)
            self._prefix = out
        return self._prefix

    def _save_rows(self, polar):
        Dim = get_sysdata().pysimbox.dim
        prefixes = self._prefixes(Dim, polar)
        cols = [field.formatColumn(index, len(prefixes)) for field, index in self._columns(Dim)]
        lines = []
        for k, pre in enumerate(prefixes):
            lines.append(pre + "".join(c[k] for c in cols) + "\n")
        self.file.write_shared("".join(lines))

    def save_perrank(self, frame_num):
        self._save_rows(polar=False)

    def save_perrank_polar(self, frame_num):
        self._save_rows(polar=True)

    def save(self, frame_num):
        if rank == 0:
            self.file.write_shared("{:d} \n".format(self.solver.gridsize))
            self.file.write_shared("#Frame {:d}  \n".format(frame_num))
        comm.barrier()
        if self.solver.coordmode == "cartesian":
            # one rank: the "ordered" mode of the reference writes the same bytes
            self.save_perrank(frame_num)
        elif self.solver.coordmode == "polar":
            self.save_perrank_polar(frame_num)
