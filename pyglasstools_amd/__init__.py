"""pyglasstools_amd: MI355X-native Irving-Kirkwood stress-field engine behind
the pyglasstools Python API.

This module mirrors pyglasstools/__init__.py (frame loop, global registries,
checkpoint-aware `analyze`); the `pyglasstools` package at the repository root
is an alias of this one, so the reference's driver scripts import it
unchanged.  All field arithmetic runs on the GPU (libpgt_ik.so); see
pyglasstools_amd/csrc/.
"""
import signal
import sys

from tqdm import tqdm

from ._native import _pyglasstools, device_count  # noqa: F401  (fails loudly when not built)

# A single-rank communicator: the engine needs no MPI (pyglasstools/__init__.py:16-18)
comm = _pyglasstools.Communicator()
rank = comm.getRank()
size = comm.getSizeGlobal()

savemode = "unordered"


def set_savemode(name):
    # As in the reference (__init__.py:21-23) this binds a local name only.
    savemode = name  # noqa: F841


_default_excepthook = sys.excepthook


def _pyglasstools_sys_excepthook(type, value, traceback):
    """Print the exception, flush, then abort the communicator (a no-op on one rank)."""
    _default_excepthook(type, value, traceback)
    sys.stdout.flush()
    sys.stderr.flush()
    comm.abort(1)


sys.excepthook = _pyglasstools_sys_excepthook

loggers_list = []
solvers_list = []
globalsysdata = None
globalpotential = None


def set_sysdata(sysdata):
    global globalsysdata
    globalsysdata = sysdata


def update_sysdata(frame_num):
    globalsysdata.update(frame_num)


def get_sysdata():
    return globalsysdata


def set_potential(potential):
    global globalpotential
    globalpotential = potential


def get_potential():
    return globalpotential


def reset():
    """Forget every registered solver and logger (pyglasstools/__init__.py:66-70)."""
    loggers_list.clear()
    solvers_list.clear()


import atexit  # noqa: E402

atexit.register(reset)


class DelayedInterrupt(object):
    """Hold SIGINT/SIGTERM until the block ends, so a frame's output and its
    checkpoint are written together (pyglasstools/__init__.py:88-102)."""

    def __enter__(self):
        self.signal_received = False
        self.old_handler = {}
        self.old_handler[signal.SIGINT] = signal.signal(signal.SIGINT, self.handler)
        self.old_handler[signal.SIGTERM] = signal.signal(signal.SIGTERM, self.handler)

    def handler(self, sig, frame):
        self.signal_received = (sig, frame)
        print(f"Signal {sig} received by Process [{rank}]. Delaying KeyboardInterrupt.")

    def __exit__(self, type, value, traceback):
        # restore each handler and replay the held signal through it, SIGINT's
        # first (its default handler raises KeyboardInterrupt)
        for sig in [signal.SIGINT, signal.SIGTERM]:
            signal.signal(sig, self.old_handler[sig])
            if self.signal_received:
                handler = self.old_handler[sig]
                if callable(handler):
                    handler(*self.signal_received)
                elif handler == signal.SIG_DFL and sig == self.signal_received[0]:
                    signal.raise_signal(sig)


def analyze(frame_list):
    """Run every solver on every frame, then save every logger and the checkpoint.

    Mirrors pyglasstools/__init__.py:107-138: update -> run for each solver,
    then (with interrupts held) each logger's save, the checkpoint line
    "Frame {n+1}" and a barrier.
    """
    totallength = len(frame_list)
    frame_list = globalsysdata.setup_checkpoint(frame_list)
    if not frame_list:
        if rank == 0:
            print("No more frames to analyze. Check your logfiles/output")
        return
    if rank == 0:
        progressbar = tqdm(total=totallength, file=sys.stdout, leave=False, initial=frame_list[0])
        print("")
    for frame_num in frame_list:
        for solver in solvers_list:
            solver.update(frame_num)
            solver.run()
        with DelayedInterrupt():
            for logger in loggers_list:
                logger.save(frame_num)
            if rank == 0 and globalsysdata.checkpointfile is not None:
                with open(globalsysdata.checkpointfile, "w") as f:
                    f.write("Frame {}".format(frame_num + 1))
            if rank == 0:
                progressbar.update(1)
                print("")
            comm.barrier()
    if rank == 0:
        progressbar.close()


from . import utils  # noqa: E402,F401
from . import potential  # noqa: E402,F401
from . import irvingkirkwood  # noqa: E402,F401
from .irvingkirkwood import cgfunc  # noqa: E402,F401
from . import io  # noqa: E402,F401
