"""Simulation box wrapper and GSD trajectory ingest (mirrors pyglasstools/utils.py)."""
import os

import numpy as np

from . import _pyglasstools, gsd_reader
from . import set_sysdata, rank


class simbox(object):
    R"""Simulation box; attribute access goes to the C++ SimBox (utils.py:7-39)."""

    def __init__(self, Lx=1.0, Ly=1.0, Lz=1.0, origin=np.zeros(3).astype("float64"), ndim=3, L=None, volume=None):
        if ndim == 2:
            Lz = 1.0
        if L is not None:
            self.cppsimbox = _pyglasstools.SimBox(L, origin, ndim)
        else:
            boxsize = np.array([Lx, Ly, Lz]).astype("float64")
            self.cppsimbox = _pyglasstools.SimBox(boxsize, origin, ndim)

    def __getattr__(self, attr):
        if attr == "cppsimbox":
            raise AttributeError(attr)
        return getattr(self.cppsimbox, attr)


def read_gsd(filename, checkpointfile=None, frame_num=0, ndim=2):
    """Open a HOOMD GSD trajectory and register it as the global system (utils.py:41-43)."""
    traj_gsd = gsd_reader.open(name=filename, mode="rb")
    return data_gsd(traj_gsd, checkpointfile, frame_num, ndim)


class data_gsd(object):
    R"""Particle data of one trajectory frame (utils.py:45-141).

    As in the reference, `update` rebuilds the particle system for a new
    frame but keeps the box of the frame the object was created with
    (utils.py:88-99 build `self.simbox` and pass `self.pysimbox`)."""

    def __init__(self, traj_gsd, checkpointfile, frame_num, ndim=2):
        self.frame_num = frame_num
        self.checkpointfile = checkpointfile
        self.traj = traj_gsd
        self.ndim = ndim
        frame = self.traj[frame_num]
        self.pysimbox = simbox(Lx=frame.configuration.box[0], Ly=frame.configuration.box[1],
                               Lz=frame.configuration.box[2], ndim=ndim)
        self.cppparticledata = _pyglasstools.ParticleSystem(
            self.pysimbox.cppsimbox, len(frame.particles.diameter), frame.particles.diameter,
            frame.particles.mass, frame.particles.position, frame.particles.velocity)
        set_sysdata(self)

    def setup_checkpoint(self, frame_list):
        """Create the checkpoint file, or resume from the frame it names (utils.py:71-83)."""
        if self.checkpointfile is not None and not os.path.exists(self.checkpointfile):
            if rank == 0:
                with open(self.checkpointfile, "w") as f:
                    f.write("Frame {}".format(frame_list[0]))
        elif self.checkpointfile is not None and os.path.exists(self.checkpointfile):
            with open(self.checkpointfile, "r") as f:
                frame_num = int(f.readline().split(" ")[1])
                frame_list = [i for i in range(frame_num, frame_list[-1] + 1)]
        return frame_list

    def update(self, frame_num):
        if frame_num != self.frame_num:
            frame = self.traj[frame_num]
            self.simbox = simbox(Lx=frame.configuration.box[0], Ly=frame.configuration.box[1],
                                 Lz=frame.configuration.box[2], ndim=self.ndim)
            self.cppparticledata = _pyglasstools.ParticleSystem(
                self.pysimbox.cppsimbox, len(frame.particles.diameter), frame.particles.diameter,
                frame.particles.mass, frame.particles.position, frame.particles.velocity)
            self.frame_num = frame_num

    def move_particles(self, displacement=None):
        if displacement is None:
            self.cppparticledata.moveParticles()
        else:
            self.cppparticledata.moveParticles(displacement)

    def set_displacement(self, displacement):
        self.cppparticledata.setDisplacement(displacement)

    def get_displacement(self):
        return self.cppparticledata.getDisplacement()

    def set_diameters(self, diameter):
        self.cppparticledata.setDiameter(diameter.astype("float64"))

    def get_diameters(self):
        return self.cppparticledata.getDiameter()

    def set_mass(self, mass):
        self.cppparticledata.setMass(mass.astype("float64"))

    def get_mass(self):
        return self.cppparticledata.getMass()

    def set_position(self, position):
        self.cppparticledata.setAtomPosition(position.astype("float64"))

    def get_position(self):
        return self.cppparticledata.getAtomPosition()

    def set_velocity(self, velocity):
        self.cppparticledata.setAtomVelocity(velocity.astype("float64"))

    def get_velocity(self):
        return self.cppparticledata.getAtomVelocity()

    def get_neighbors(self, point, radius):
        return self.cppparticledata.getNeighbors(point, radius)

    def get_neighborsdistance(self, tag, radius):
        return self.cppparticledata.getNeighborsDistance(tag, radius)
