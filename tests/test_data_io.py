"""Snapshot output of the facade: HDF5 layout of abie/data_io.py, the
recorder, event files, configuration files and snapshot serialization.

The files are read back twice: through ``abie_amd.h5file`` and through the
HDF5 distribution's own ``h5dump`` (an independent reader), whose raw
little-endian dump of each dataset must equal the bytes written.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from abie_amd import h5file
from abie_amd.data_io import DataIO, STEP_FIELDS
from abie_amd.recorder import SimulationDataRecorder

pytestmark = pytest.mark.skipif(not h5file.available(), reason="no libhdf5 on this machine")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _h5dump_raw(path, dset, tmp_path):
    out = tmp_path / ("dump_%s.bin" % dset.strip("/").replace("/", "_").replace("#", ""))
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", str(out), str(path)], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return out.read_bytes()


def _state(rng, n):
    return rng.normal(size=3 * n), rng.normal(size=3 * n)


def test_h5file_round_trip_and_independent_reader(tmp_path):
    path = tmp_path / "t.hdf5"
    rng = np.random.default_rng(3)
    x = rng.normal(size=(5, 7))
    h = np.arange(12, dtype=np.int64).reshape(3, 4) * 1_000_000_007
    with h5file.File(str(path), "w") as f:
        f.set_attr("G", 39.47841760435743)
        f.set_attr("n", 7)
        f.create_group("Step#0")
        f.create_dataset("Step#0/x", x)
        f.create_dataset("Step#0/hash", h)
        f.create_dataset("Step#0/none", np.zeros((0, 7)))
        f.create_dataset("loose", np.array([1.5, -2.0]))
    with h5file.File(str(path), "r") as f:
        assert f.get_attr("G") == 39.47841760435743
        assert f.get_attr("n") == 7 and isinstance(f.get_attr("n"), np.int64)
        assert f.keys() == ["Step#0", "loose"]
        assert f.keys("Step#0") == ["hash", "none", "x"]
        assert f.is_group("Step#0") and not f.is_group("loose")
        assert "Step#0/x" in f and "Step#1" not in f and "Step#0/y" not in f
        assert np.array_equal(f.read("Step#0/x"), x)
        assert np.array_equal(f.read("Step#0/hash"), h) and f.read("Step#0/hash").dtype == np.int64
        assert f.read("Step#0/none").shape == (0, 7)
    with pytest.raises(h5file.H5Error):
        h5file.File(str(tmp_path / "missing.hdf5"), "r")
    if H5DUMP is None:
        pytest.skip("h5dump not installed")
    assert _h5dump_raw(path, "/Step#0/x", tmp_path) == x.astype("<f8").tobytes()
    assert _h5dump_raw(path, "/Step#0/hash", tmp_path) == h.astype("<i8").tobytes()
    header = subprocess.run([H5DUMP, "-H", str(path)], check=True, capture_output=True, text=True).stdout
    assert 'ATTRIBUTE "G"' in header and "H5T_IEEE_F64LE" in header and "H5T_STD_I64LE" in header


def test_data_io_layout_matches_reference(tmp_path):
    """Rows, flush groups, skipped t = 0, NaN for unsupplied fields."""
    path = tmp_path / "sim.hdf5"
    io = DataIO(buf_len=3, output_file_name=str(path), collision_output_file_name=None,
                close_encounter_output_file_name=None, CONST_G=4 * np.pi ** 2)
    io.initialize_buffer(4)
    rng = np.random.default_rng(0)
    rows = []
    for k, t in enumerate([0.0, 0.5, 1.0, 1.5, 2.0, 2.0, 2.5]):
        pos, vel = _state(rng, 4)
        m = np.array([1.0, 1e-3, 3e-4, 0.0])
        kw = dict(radii=np.full(4, 0.01), names=np.arange(4) + 100, ptypes=np.zeros(4), a=rng.random(4),
                  e=rng.random(4), i=rng.random(4), energy=-0.1 - k)
        if t == 1.5:
            kw = {}  # only the mandatory fields
        io.store_state(t, pos, vel, m, **kw)
        if t != 0.0 and (not rows or rows[-1][0] != t):
            rows.append((t, pos.reshape(-1, 3), vel.reshape(-1, 3), kw))
    io.close()
    assert [r[0] for r in rows] == [0.5, 1.0, 1.5, 2.0, 2.5]
    with h5file.File(str(path), "r") as f:
        assert f.get_attr("G") == 4 * np.pi ** 2
        assert f.keys() == ["Step#0", "Step#1"]
        for g, part in (("Step#0", rows[:3]), ("Step#1", rows[3:])):
            assert sorted(f.keys(g)) == sorted(STEP_FIELDS)
            assert np.array_equal(f.read(g + "/time"), [r[0] for r in part])
            for c, key in enumerate("xyz"):
                assert np.array_equal(f.read("%s/%s" % (g, key)), np.array([r[1][:, c] for r in part]))
                assert np.array_equal(f.read("%s/v%s" % (g, key)), np.array([r[2][:, c] for r in part]))
            for key in STEP_FIELDS:
                d = f.read("%s/%s" % (g, key))
                assert d.dtype == np.float64
                assert d.shape == ((len(part),) if key in ("time", "energy") else (len(part), 4))
        hsh = f.read("Step#0/hash")
        assert np.array_equal(hsh[0], np.arange(4) + 100)
        assert np.isnan(hsh[2]).all() and np.isnan(f.read("Step#0/energy")[2])
        assert np.array_equal(f.read("Step#1/energy"), [-0.1 - 4, -0.1 - 6])
    if H5DUMP is not None:
        x0 = np.array([r[1][:, 0] for r in rows[:3]])
        assert _h5dump_raw(path, "/Step#0/x", tmp_path) == x0.astype("<f8").tobytes()


def test_event_files_and_first_allocation_cleanup(tmp_path):
    ce = tmp_path / "ce.txt"
    coll = tmp_path / "coll.txt"
    ce.write_text("stale\n")
    io = DataIO(output_file_name=None, collision_output_file_name=str(coll), close_encounter_output_file_name=str(ce))
    io.initialize_buffer(3)
    assert not ce.exists()
    io.store_close_encounters(np.array([1.25, 0, 2, 1e-3, 2.5, 1, 2, 2e-3]))
    io.store_collisions(np.array([3.0, 1, 2, 4e-5]))
    io.reset_buffer()
    io.initialize_buffer(2)  # re-allocation after a merge keeps the event files
    assert coll.exists() and ce.exists()
    lines = ce.read_text().splitlines()
    assert lines == ["# Time, Particle 1, Particle 2, Distance", "1.25, 0, 2, 0.001", "2.5, 1, 2, 0.002"]
    back = np.loadtxt(str(coll), delimiter=",")
    assert np.array_equal(back, [3.0, 1, 2, 4e-5])


def test_recorder_keeps_columns_across_a_merge(tmp_path):
    io = DataIO(buf_len=2, output_file_name=str(tmp_path / "r.hdf5"), collision_output_file_name=None,
                close_encounter_output_file_name=None)
    rec = io.recorder
    rec.set_monitored_quantities(["x", "time", "mass"])
    rec.start()
    io.initialize_buffer(3)
    names = np.array([11, 22, 33])
    for t in (1.0, 2.0):
        io.store_state(t, np.arange(9.0) * t, np.zeros(9), np.array([1.0, 2.0, 3.0]), names=names)
    io.flush()
    # body 22 merged away: 2 bodies remain, hashes 11 and 33
    io.reset_buffer()
    io.initialize_buffer(2)
    io.store_state(3.0, np.array([1.0, 0, 0, 7.0, 0, 0]), np.zeros(6), np.array([1.5, 3.0]), names=names[[0, 2]])
    io.close()
    x = rec.data["x"]
    assert x.shape == (3, 3)
    assert np.array_equal(x[:2], [[0, 3, 6], [0, 6, 12]])
    assert x[2, 0] == 1.0 and np.isnan(x[2, 1]) and x[2, 2] == 7.0
    assert np.array_equal(rec.data["time"], [1.0, 2.0, 3.0])
    assert "a" not in rec.data
    sel = SimulationDataRecorder(particles=[2], quantities=["mass"])
    sel.start()
    sel.record({"hash": np.array([[5.0, 6.0, 7.0]]), "mass": np.array([[1.0, 2.0, 3.0]])})
    assert np.array_equal(sel.data["mass"], [[3.0]])
    idle = SimulationDataRecorder()
    idle.record({"hash": np.array([[5.0]]), "a": np.array([[1.0]])})
    assert idle.data == {}


def test_snapshot_serialization_merges_steps(tmp_path):
    from abie_amd.snapshot_serialization import serialize
    src = tmp_path / "p.hdf5"
    io = DataIO(buf_len=2, output_file_name=str(src), collision_output_file_name=None,
                close_encounter_output_file_name=None)
    io.initialize_buffer(3)
    names = np.array([11, 22, 33])
    for t in (1.0, 2.0, 3.0):
        io.store_state(t, np.arange(9.0) + t, np.zeros(9), np.ones(3), names=names)
    io.flush()
    io.reset_buffer()
    io.initialize_buffer(2)
    io.store_state(4.0, np.array([1.0, 0, 0, 2.0, 0, 0]), np.zeros(6), np.ones(2), names=names[[0, 2]])
    io.close()
    out = tmp_path / "merged.hdf5"
    serialize(str(src), str(out))
    with h5file.File(str(out), "r") as f:
        assert np.array_equal(f.read("time"), [1.0, 2.0, 3.0, 4.0])
        x = f.read("x")
        assert np.array_equal(x[:3], [[1, 4, 7], [2, 5, 8], [3, 6, 9]])
        assert x[3, 0] == 1.0 and np.isnan(x[3, 1]) and x[3, 2] == 2.0
        hsh = f.read("hash")
        assert hsh.dtype == np.int64 and np.array_equal(hsh[3], [11, -1, 33])
    only = tmp_path / "mass_only.hdf5"
    serialize(str(src), str(only), dataset="mass")
    with h5file.File(str(only), "r") as f:
        assert f.keys() == ["mass"]


def test_config_file_round_trip(tmp_path):
    state = np.arange(54, dtype=np.float64) * 0.1
    masses = np.linspace(1.0, 1e-8, 9)
    path = tmp_path / "ss.toml"
    DataIO.toml_generator(state, 39.478, 0.0, 10.0, 0.01, masses, str(path))
    cfg = DataIO.parse_config_file(str(path))
    assert cfg["physical_params"]["G"] == 39.478
    assert cfg["integration"] == {"t0": 0.0, "tf": 10.0, "h": 0.01}
    earth = cfg["initial_conds"]["Earth"]
    assert earth == list(state[9:12]) + list(state[36:39]) + [masses[3]]  # Earth is body 3 of 9

    class Sink(object):
        def __init__(self):
            self.rows = []

        def add(self, pos, vel, mass, name):
            self.rows.append((name, list(pos), list(vel), mass))

    sink = Sink()
    DataIO.ic_populate(cfg["initial_conds"], sink, names=["Sun", "Earth"])
    assert [r[0] for r in sink.rows] == ["Sun", "Earth"]
    assert sink.rows[1][1] == list(state[9:12]) and sink.rows[1][3] == masses[3]
