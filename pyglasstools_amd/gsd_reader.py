"""Reader for HOOMD-blue GSD trajectory files (GSD container 1.x and 2.x,
"hoomd" schema), with the frame semantics of gsd.hoomd.

The reference reads trajectories through the third-party `gsd.hoomd`
package (pyglasstools/utils.py:5,41-69), which is not a dependency here.  This
module reads the same files with numpy only.

GSD container layout (little endian):
  header, 256 bytes: magic u64 (0x65DF65DF65DF65DF), index_location u64,
      index_allocated_entries u64, namelist_location u64,
      namelist_allocated_entries u64, schema_version u32, gsd_version u32,
      application char[64], schema char[64], reserved char[80]
  index: index_allocated_entries entries of 32 bytes
      frame u64, N u64, location i64, M u32, id u16, type u8, flags u8
      (location 0 marks an unused entry)
  namelist: namelist_allocated_entries * 64 bytes of NUL-terminated chunk
      names; a chunk's id is the ordinal of its name
Chunk data: N rows x M columns of `type` at `location`.

Frame semantics (as gsd.hoomd): a chunk missing from frame k is taken from
frame 0 if frame 0 has it, otherwise the schema default (mass 1, diameter 1,
zero positions/velocities, box [1,1,1,0,0,0]).
"""
import builtins
import mmap
import os

import numpy as np

GSD_MAGIC = 0x65DF65DF65DF65DF
_HEADER = np.dtype([("magic", "<u8"), ("index_location", "<u8"), ("index_allocated_entries", "<u8"),
                    ("namelist_location", "<u8"), ("namelist_allocated_entries", "<u8"),
                    ("schema_version", "<u4"), ("gsd_version", "<u4"), ("application", "S64"),
                    ("schema", "S64"), ("reserved", "S80")])
_INDEX = np.dtype([("frame", "<u8"), ("N", "<u8"), ("location", "<i8"), ("M", "<u4"), ("id", "<u2"),
                   ("type", "u1"), ("flags", "u1")])
_TYPES = {1: np.uint8, 2: np.uint16, 3: np.uint32, 4: np.uint64, 5: np.int8, 6: np.int16, 7: np.int32,
          8: np.int64, 9: np.float32, 10: np.float64, 11: np.uint8}

# per-particle chunks: (columns, dtype, default value)
_PARTICLE_CHUNKS = {
    "position": (3, np.float32, 0.0),
    "orientation": (4, np.float32, None),
    "typeid": (1, np.uint32, 0),
    "mass": (1, np.float32, 1.0),
    "charge": (1, np.float32, 0.0),
    "diameter": (1, np.float32, 1.0),
    "body": (1, np.int32, -1),
    "moment_inertia": (3, np.float32, 0.0),
    "velocity": (3, np.float32, 0.0),
    "angmom": (4, np.float32, 0.0),
    "image": (3, np.int32, 0),
}


class GSDFormatError(ValueError):
    pass


class GSDFile(object):
    """Low-level chunk access to a GSD file (read only, memory mapped)."""

    def __init__(self, name):
        self.name = name
        self._fh = builtins.open(name, "rb")
        size = os.fstat(self._fh.fileno()).st_size
        if size < _HEADER.itemsize:
            raise GSDFormatError("%s: file too short for a GSD header" % name)
        self._mm = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        h = np.frombuffer(self._mm, dtype=_HEADER, count=1, offset=0)[0]
        if int(h["magic"]) != GSD_MAGIC:
            raise GSDFormatError("%s: not a GSD file (bad magic)" % name)
        self.gsd_version = (int(h["gsd_version"]) >> 16, int(h["gsd_version"]) & 0xFFFF)
        if self.gsd_version[0] not in (1, 2):
            raise GSDFormatError("%s: unsupported GSD version %d.%d" % ((name,) + self.gsd_version))
        self.schema_version = (int(h["schema_version"]) >> 16, int(h["schema_version"]) & 0xFFFF)
        self.application = h["application"].split(b"\0")[0].decode()
        self.schema = h["schema"].split(b"\0")[0].decode()
        nl_loc, nl_n = int(h["namelist_location"]), int(h["namelist_allocated_entries"])
        raw = bytes(self._mm[nl_loc:nl_loc + 64 * nl_n])
        self.names = [s.decode() for s in raw.split(b"\0") if s]
        ix_loc, ix_n = int(h["index_location"]), int(h["index_allocated_entries"])
        if ix_loc + ix_n * _INDEX.itemsize > size:
            raise GSDFormatError("%s: index runs past the end of the file" % name)
        index = np.frombuffer(self._mm, dtype=_INDEX, count=ix_n, offset=ix_loc)
        index = index[index["location"] != 0]
        self._entries = {}
        for e in index:
            self._entries[(int(e["frame"]), int(e["id"]))] = (int(e["N"]), int(e["M"]), int(e["location"]), int(e["type"]))
        self.nframes = int(index["frame"].max()) + 1 if index.size else 0
        self._ids = {n: k for k, n in enumerate(self.names)}

    def chunk_exists(self, frame, name):
        k = self._ids.get(name)
        return k is not None and (frame, k) in self._entries

    def read_chunk(self, frame, name):
        k = self._ids.get(name)
        if k is None or (frame, k) not in self._entries:
            raise KeyError("frame %d has no chunk %s" % (frame, name))
        N, M, loc, typ = self._entries[(frame, k)]
        if typ not in _TYPES:
            raise GSDFormatError("chunk %s has unknown type %d" % (name, typ))
        dt = np.dtype(_TYPES[typ]).newbyteorder("<")
        a = np.frombuffer(self._mm, dtype=dt, count=N * M, offset=loc).copy()
        return a.reshape(N, M) if M > 1 else a

    def close(self):
        # arrays handed out are copies, so the map can close now
        if self._mm is not None:
            self._mm.close()
            self._mm = None
        if self._fh is not None:
            self._fh.close()
            self._fh = None


class ConfigurationData(object):
    def __init__(self):
        self.step = 0
        self.dimensions = 3
        self.box = np.array([1, 1, 1, 0, 0, 0], dtype=np.float32)


class ParticleData(object):
    def __init__(self):
        self.N = 0
        self.types = ["A"]
        for name in _PARTICLE_CHUNKS:
            setattr(self, name, None)


class Frame(object):
    def __init__(self):
        self.configuration = ConfigurationData()
        self.particles = ParticleData()


class HOOMDTrajectory(object):
    """Sequence of frames with gsd.hoomd's `traj[i].configuration/.particles` access."""

    def __init__(self, gsdfile):
        self.file = gsdfile
        if gsdfile.schema != "hoomd":
            raise GSDFormatError("%s: schema %r is not 'hoomd'" % (gsdfile.name, gsdfile.schema))

    def __len__(self):
        return self.file.nframes

    def __iter__(self):
        for k in range(len(self)):
            yield self[k]

    def __getitem__(self, key):
        if isinstance(key, slice):
            return [self[k] for k in range(*key.indices(len(self)))]
        k = int(key)
        if k < 0:
            k += len(self)
        if not 0 <= k < len(self):
            raise IndexError("frame %d out of range" % key)
        return self._read_frame(k)

    def _chunk(self, frame, name):
        f = self.file
        if f.chunk_exists(frame, name):
            return f.read_chunk(frame, name)
        if frame != 0 and f.chunk_exists(0, name):
            return f.read_chunk(0, name)
        return None

    def _read_frame(self, k):
        fr = Frame()
        c = fr.configuration
        step = self._chunk(k, "configuration/step")
        if step is not None:
            c.step = int(step[0])
        box = self._chunk(k, "configuration/box")
        if box is not None:
            c.box = box.astype(np.float32)
        dims = self._chunk(k, "configuration/dimensions")
        c.dimensions = int(dims[0]) if dims is not None else (2 if float(c.box[2]) == 0.0 else 3)
        p = fr.particles
        n = self._chunk(k, "particles/N")
        p.N = int(n[0]) if n is not None else 0
        types = self._chunk(k, "particles/types")
        if types is not None:
            t = types.reshape(-1, types.shape[-1]) if types.ndim > 1 else types.reshape(1, -1)
            p.types = [bytes(row).split(b"\0")[0].decode() for row in t]
        for name, (cols, dt, default) in _PARTICLE_CHUNKS.items():
            a = self._chunk(k, "particles/" + name)
            if a is None:
                if name == "orientation":
                    a = np.zeros((p.N, 4), dtype=dt)
                    a[:, 0] = 1.0
                else:
                    shape = (p.N, cols) if cols > 1 else (p.N,)
                    a = np.full(shape, default, dtype=dt)
            elif a.shape[0] != p.N:
                raise GSDFormatError("frame %d: particles/%s has %d rows for N=%d" % (k, name, a.shape[0], p.N))
            setattr(p, name, a)
        return fr

    def close(self):
        self.file.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def open(name, mode="rb"):  # noqa: A001  (mirrors gsd.hoomd.open)
    if mode not in ("rb", "r"):
        raise ValueError("only read mode is supported")
    return HOOMDTrajectory(GSDFile(name))
