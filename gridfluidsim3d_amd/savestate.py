"""Reader (and writer) for GridFluidSim3D's binary save states, as scenes
for the transfer + advection step.

Layout written by FluidSimulationSaveState::saveState
(fluidsimulationsavestate.cpp:31-99), native byte order:

    int i, j, k; double dx; int frame; int n_marker; int n_diffuse;
    int n_solid; bool brick_grid (1 byte)
    float marker positions  [n_marker * 3]   (x1, y1, z1, x2, ...)
    float marker velocities [n_marker * 3]
    float diffuse positions [n_diffuse * 3]
    float diffuse velocities[n_diffuse * 3]
    float diffuse lifetimes [n_diffuse]
    char  diffuse types     [n_diffuse]
    int   solid cells       [n_solid * 3]    (i1, j1, k1, i2, ...)
    (the FluidBrickGrid save state, when brick_grid is set)

The block offsets follow loadState (fluidsimulationsavestate.cpp:106-146).
``SaveStateReader`` keeps the method names of pyfluid's
FluidSimulationSaveState (pyfluid/fluidsimulationsavestate.py:18-200) but
returns numpy arrays, memory-mapped from the file, so a 40M-particle state is
not read into memory until it is used.  ``load_state`` returns False for a
file that cannot be opened or whose header is short, as loadState does.
"""
import os
import struct

import numpy as np

HEADER = struct.Struct("=iiidiiii?")   # 37 bytes, no padding


class SaveStateReader:
    def __init__(self):
        self._path = None
        self._hdr = None

    # ------------------------------------------------------------ loading

    def load_state(self, filename):
        self.close_state()
        try:
            size = os.path.getsize(filename)
            with open(filename, "rb") as f:
                raw = f.read(HEADER.size)
        except OSError:
            return False
        if len(raw) < HEADER.size:
            return False
        i, j, k, dx, frame, nm, nd, ns, brick = HEADER.unpack(raw)
        if min(i, j, k, nm, nd, ns) < 0:
            return False
        off = {}
        off["mp_pos"] = HEADER.size
        off["mp_vel"] = off["mp_pos"] + 12 * nm
        off["dp_pos"] = off["mp_vel"] + 12 * nm
        off["dp_vel"] = off["dp_pos"] + 12 * nd
        off["dp_life"] = off["dp_vel"] + 12 * nd
        off["dp_type"] = off["dp_life"] + 4 * nd
        off["solid"] = off["dp_type"] + nd
        off["end"] = off["solid"] + 12 * ns
        if off["end"] > size:
            return False
        self._path = filename
        self._hdr = dict(dims=(i, j, k), dx=dx, frame=frame, nm=nm, nd=nd, ns=ns,
                         brick=bool(brick), off=off)
        return True

    def close_state(self):
        self._path = None
        self._hdr = None

    def is_load_state_initialized(self):
        return self._hdr is not None

    def _need(self):
        if self._hdr is None:
            raise RuntimeError("save state not loaded")
        return self._hdr

    # ------------------------------------------------------------ header

    def get_grid_dimensions(self):
        return self._need()["dims"]

    def get_cell_size(self):
        return self._need()["dx"]

    def get_current_frame(self):
        return self._need()["frame"]

    def get_num_marker_particles(self):
        return self._need()["nm"]

    def get_num_diffuse_particles(self):
        return self._need()["nd"]

    def get_num_solid_cells(self):
        return self._need()["ns"]

    def is_fluid_brick_grid_enabled(self):
        return self._need()["brick"]

    # ------------------------------------------------------------ blocks

    def _block(self, key, count, dtype, width, startidx, endidx):
        h = self._need()
        startidx = 0 if startidx is None else int(startidx)
        endidx = count if endidx is None else int(endidx)
        if not (0 <= startidx <= endidx <= count):
            raise IndexError("range [%d, %d) outside [0, %d)" % (startidx, endidx, count))
        n = endidx - startidx
        item = np.dtype(dtype).itemsize * width
        if n == 0:
            return np.zeros((0, width) if width > 1 else (0,), dtype)
        mm = np.memmap(self._path, dtype=dtype, mode="r", offset=h["off"][key] + startidx * item,
                       shape=(n, width) if width > 1 else (n,))
        return mm

    def get_marker_particle_positions(self, startidx=None, endidx=None):
        return self._block("mp_pos", self._need()["nm"], np.float32, 3, startidx, endidx)

    def get_marker_particle_velocities(self, startidx=None, endidx=None):
        return self._block("mp_vel", self._need()["nm"], np.float32, 3, startidx, endidx)

    def get_diffuse_particle_positions(self, startidx=None, endidx=None):
        return self._block("dp_pos", self._need()["nd"], np.float32, 3, startidx, endidx)

    def get_diffuse_particle_velocities(self, startidx=None, endidx=None):
        return self._block("dp_vel", self._need()["nd"], np.float32, 3, startidx, endidx)

    def get_diffuse_particle_lifetimes(self, startidx=None, endidx=None):
        return self._block("dp_life", self._need()["nd"], np.float32, 1, startidx, endidx)

    def get_diffuse_particle_types(self, startidx=None, endidx=None):
        return self._block("dp_type", self._need()["nd"], np.int8, 1, startidx, endidx)

    def get_solid_cells(self, startidx=None, endidx=None):
        return self._block("solid", self._need()["ns"], np.int32, 3, startidx, endidx)

    # ------------------------------------------------------------ scenes

    def to_scene(self, dt, body_force=(0.0, -9.8, 0.0), name=None):
        """The marker particles as a scene for the plain step (positions and
        velocities copied into contiguous float32 arrays)."""
        from .scenes import SceneSpec
        h = self._need()
        pos = np.array(self.get_marker_particle_positions(), np.float32)
        vel = np.array(self.get_marker_particle_velocities(), np.float32)
        return SceneSpec(name or os.path.basename(self._path), h["dims"], h["dx"], pos, vel, dt,
                         body_force)


def write_state(filename, dims, dx, marker_pos, marker_vel, frame=0, diffuse_pos=None,
                diffuse_vel=None, diffuse_life=None, diffuse_type=None, solid_cells=None):
    """Write a save state in the layout above (no FluidBrickGrid block):
    the input side of FluidSimulationSaveState::saveState, for building
    scenes the reference's loader and this reader both read."""
    mp = np.ascontiguousarray(marker_pos, np.float32).reshape(-1, 3)
    mv = np.ascontiguousarray(marker_vel, np.float32).reshape(-1, 3)
    if mp.shape != mv.shape:
        raise ValueError("marker positions and velocities differ in length")
    nd = 0 if diffuse_pos is None else len(diffuse_pos)
    dp = np.ascontiguousarray(diffuse_pos if nd else np.zeros((0, 3)), np.float32).reshape(-1, 3)
    dv = np.ascontiguousarray(diffuse_vel if nd else np.zeros((0, 3)), np.float32).reshape(-1, 3)
    dl = np.ascontiguousarray(diffuse_life if nd else np.zeros(0), np.float32).reshape(-1)
    dtp = np.ascontiguousarray(diffuse_type if nd else np.zeros(0), np.int8).reshape(-1)
    if not (len(dp) == len(dv) == len(dl) == len(dtp) == nd):
        raise ValueError("diffuse particle arrays differ in length")
    sc = np.ascontiguousarray(solid_cells if solid_cells is not None else np.zeros((0, 3)),
                              np.int32).reshape(-1, 3)
    i, j, k = (int(d) for d in dims)
    with open(filename, "wb") as f:
        f.write(HEADER.pack(i, j, k, float(dx), int(frame), len(mp), nd, len(sc), False))
        for a in (mp, mv, dp, dv, dl, dtp, sc):
            f.write(a.tobytes())
