"""The save-state reader (gridfluidsim3d_amd/savestate.py) against the byte
layout FluidSimulationSaveState::saveState writes
(fluidsimulationsavestate.cpp:31-99; block offsets of loadState, :106-146).
The reference's save-state class cannot be compiled here (it includes the
whole FluidSimulation and its OpenCL headers), so the layout is pinned by
building the bytes field by field in the writer's statement order."""
import struct

import numpy as np
import pytest

from gridfluidsim3d_amd import savestate


def reference_bytes(dims, dx, frame, mp, mv, dp, dv, dl, dt, sc, brick=False, tail=b""):
    """The file saveState writes, one _write* call at a time."""
    b = b""
    for v in dims:
        b += struct.pack("=i", v)          # _writeInt(&i), (&j), (&k)
    b += struct.pack("=d", dx)             # _writeDouble(&dx)
    b += struct.pack("=i", frame)          # current frame
    b += struct.pack("=i", len(mp))        # marker particles
    b += struct.pack("=i", len(dp))        # diffuse particles
    b += struct.pack("=i", len(sc))        # solid cells
    b += struct.pack("=?", brick)          # _writeBool (sizeof(bool) == 1)
    for x in mp:
        b += struct.pack("=3f", *x)
    for x in mv:
        b += struct.pack("=3f", *x)
    for x in dp:
        b += struct.pack("=3f", *x)
    for x in dv:
        b += struct.pack("=3f", *x)
    for x in dl:
        b += struct.pack("=f", x)
    for x in dt:
        b += struct.pack("=b", x)
    for x in sc:
        b += struct.pack("=3i", *x)
    return b + tail


def sample(seed=5, nm=37, nd=11, ns=6):
    rng = np.random.default_rng(seed)
    mp = rng.uniform(0, 2, size=(nm, 3)).astype(np.float32)
    mv = rng.normal(size=(nm, 3)).astype(np.float32)
    dp = rng.uniform(0, 2, size=(nd, 3)).astype(np.float32)
    dv = rng.normal(size=(nd, 3)).astype(np.float32)
    dl = rng.uniform(0, 5, size=nd).astype(np.float32)
    dt = rng.integers(0, 3, size=nd).astype(np.int8)
    sc = rng.integers(0, 16, size=(ns, 3)).astype(np.int32)
    return mp, mv, dp, dv, dl, dt, sc


def test_writer_matches_reference_layout(tmp_path):
    mp, mv, dp, dv, dl, dt, sc = sample()
    f = tmp_path / "a.state"
    savestate.write_state(str(f), (16, 12, 8), 0.125, mp, mv, frame=42, diffuse_pos=dp,
                          diffuse_vel=dv, diffuse_life=dl, diffuse_type=dt, solid_cells=sc)
    assert f.read_bytes() == reference_bytes((16, 12, 8), 0.125, 42, mp, mv, dp, dv, dl, dt, sc)
    assert savestate.HEADER.size == 37


def test_reader_reads_reference_bytes(tmp_path):
    mp, mv, dp, dv, dl, dt, sc = sample(6)
    f = tmp_path / "b.state"
    # a brick-grid block after the solid cells is skipped like loadState does
    f.write_bytes(reference_bytes((20, 10, 5), 0.05, 7, mp, mv, dp, dv, dl, dt, sc, brick=True,
                                  tail=b"\x01" * 64))
    r = savestate.SaveStateReader()
    assert r.load_state(str(f))
    assert r.get_grid_dimensions() == (20, 10, 5)
    assert r.get_cell_size() == 0.05
    assert r.get_current_frame() == 7
    assert r.get_num_marker_particles() == len(mp)
    assert r.get_num_diffuse_particles() == len(dp)
    assert r.get_num_solid_cells() == len(sc)
    assert r.is_fluid_brick_grid_enabled()
    assert np.array_equal(r.get_marker_particle_positions(), mp)
    assert np.array_equal(r.get_marker_particle_velocities(), mv)
    assert np.array_equal(r.get_diffuse_particle_positions(), dp)
    assert np.array_equal(r.get_diffuse_particle_velocities(), dv)
    assert np.array_equal(r.get_diffuse_particle_lifetimes(), dl)
    assert np.array_equal(r.get_diffuse_particle_types(), dt)
    assert np.array_equal(r.get_solid_cells(), sc)
    assert np.array_equal(r.get_marker_particle_positions(5, 9), mp[5:9])
    assert np.array_equal(r.get_marker_particle_velocities(36, 37), mv[36:37])
    assert r.get_solid_cells(2, 2).shape == (0, 3)
    with pytest.raises(IndexError):
        r.get_marker_particle_positions(3, 38)
    sc_ = r.to_scene(1.0 / 120.0)
    assert sc_.dims == (20, 10, 5) and sc_.n == len(mp) and sc_.dx == 0.05
    assert np.array_equal(sc_.pos, mp) and np.array_equal(sc_.vel, mv)
    r.close_state()
    assert not r.is_load_state_initialized()


def test_reader_rejects_short_files(tmp_path):
    r = savestate.SaveStateReader()
    assert not r.load_state(str(tmp_path / "missing.state"))
    mp, mv, dp, dv, dl, dt, sc = sample(7)
    full = reference_bytes((8, 8, 8), 0.1, 0, mp, mv, dp, dv, dl, dt, sc)
    f = tmp_path / "c.state"
    f.write_bytes(full[:30])                   # header cut short: loadState fails
    assert not r.load_state(str(f))
    f.write_bytes(full[:-5])                   # a block cut short
    assert not r.load_state(str(f))
    f.write_bytes(full)
    assert r.load_state(str(f))
    with pytest.raises(RuntimeError):
        savestate.SaveStateReader().get_cell_size()


def test_empty_state(tmp_path):
    f = tmp_path / "d.state"
    savestate.write_state(str(f), (4, 4, 4), 0.25, np.zeros((0, 3)), np.zeros((0, 3)))
    r = savestate.SaveStateReader()
    assert r.load_state(str(f))
    assert r.get_marker_particle_positions().shape == (0, 3)
    assert r.to_scene(0.01).n == 0
