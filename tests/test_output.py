"""Output files (version-1 .npz) and their reader, on the CPU
(solver/solver.py:605-747, solver/utils.py:302-341)."""
import collections
import io
import os
import pickle
import types
import warnings
import zipfile

import numpy as np
import pytest

from pysph_amd.base import api as base
from pysph_amd.solver import utils
from pysph_amd.solver.solver import Solver


def _arrays():
    rng = np.random.default_rng(7)
    fl = base.get_particle_array(name="fluid", type=base.Fluid, x=rng.random(11),
                                 y=rng.random(11), rho=1000 + rng.random(11), device="cpu",
                                 c0=62.5)
    wl = base.get_particle_array(name="wall", type=base.Solid, x=rng.random(4), device="cpu")
    fl.add_property({"name": "_x0", "type": "double", "default": 3.0})
    return fl, wl


def _solver_with(arrays, directory, fname="drop"):
    s = Solver(dim=2)
    s.particles = types.SimpleNamespace(arrays=list(arrays))
    s.set_output_directory(str(directory))
    s.set_output_fname(fname)
    return s


def test_dump_layout_matches_version_1(tmp_path):
    """Three members; version 1; arrays = {name: {prop: array, constants}};
    solver_data = {dt, t, count}; file <fname>_<count>.npz (solver.py:647-699)."""
    fl, wl = _arrays()
    s = _solver_with((fl, wl), tmp_path / "out")
    s.t, s.count = 0.125, 40
    path = s.dump_output(2.5e-5)
    s.wait_output()
    assert os.path.basename(path) == "drop_40.npz"
    with zipfile.ZipFile(path) as z:
        assert sorted(z.namelist()) == ["arrays.npy", "solver_data.npy", "version.npy"]
        assert all(i.compress_type == zipfile.ZIP_STORED for i in z.infolist())
    # a file this code wrote: numpy's own pickle-enabled loader reads the same data
    with np.load(path, allow_pickle=True) as raw:
        assert int(raw["version"]) == 1
        arrays = raw["arrays"][()]
        sd = raw["solver_data"][()]
    assert sd == {"dt": 2.5e-5, "t": 0.125, "count": 40}
    assert set(arrays) == {"fluid", "wall"}
    # non-detailed output: no '_' properties except _tmpx/_tmpy/_tmpz
    assert "_x0" not in arrays["fluid"] and "_tmpx" in arrays["fluid"]
    assert arrays["fluid"]["c0"] == 62.5
    assert np.array_equal(arrays["fluid"]["rho"], fl.get("rho"))
    assert arrays["fluid"]["tag"].dtype == np.int64 and arrays["fluid"]["local"].dtype == np.int32
    got = utils.read_output(path)
    for name in arrays:
        assert set(got["arrays"][name]) == set(arrays[name])
        for k, v in arrays[name].items():
            g = got["arrays"][name][k]
            if isinstance(v, np.ndarray):
                assert g.dtype == v.dtype and np.array_equal(g, v), (name, k)
            else:
                assert g == v
    assert got["solver_data"] == sd


def test_detailed_output_has_every_property(tmp_path):
    fl, wl = _arrays()
    s = _solver_with((fl, wl), tmp_path)
    s.set_output_printing_level(True)
    s.dump_output(1e-5)
    s.wait_output()
    got = utils.read_output(tmp_path / "drop_0.npz")
    assert np.array_equal(got["arrays"]["fluid"]["_x0"], np.full(11, 3.0))
    assert set(fl.properties) <= set(got["arrays"]["fluid"])


def test_load_builds_particle_arrays(tmp_path):
    fl, wl = _arrays()
    s = _solver_with((fl, wl), tmp_path)
    s.dump_output(1e-5)
    s.wait_output()
    out = utils.load(tmp_path / "drop_0.npz", device="cpu")
    f2 = out["arrays"]["fluid"]
    assert f2.get_number_of_particles() == 11 and f2.num_real_particles == 11
    for k in ("x", "y", "rho", "tag", "idx"):
        assert np.array_equal(f2.get(k), fl.get(k)), k
    assert f2.get_type_name("local") == "int" and f2.get_type_name("tag") == "long"
    assert f2.c0 == 62.5
    # the particle type is not in the file (the reference's loader has the same gap)
    assert out["arrays"]["wall"].particle_type == base.Fluid
    assert out["solver_data"]["count"] == 0


def test_output_counts_and_missing_count(tmp_path):
    fl, wl = _arrays()
    s = _solver_with((fl, wl), tmp_path)
    for c in (0, 100, 20, 3):
        s.count = c
        s.dump_output(1e-5)
    s.wait_output()
    (tmp_path / "drop_notes.npz").write_bytes(b"")
    (tmp_path / "other_7.npz").write_bytes(b"")
    assert s.load_output('?') == ["0", "3", "20", "100"]
    with pytest.raises(IOError, match=r"Valid iteration counts are \['0', '3', '20', '100'\]"):
        s.load_output(5)


def _npz_with_arrays_object(path, obj, protocol=3):
    """An output file whose 'arrays' member pickles ``obj`` (the test's own file)."""
    buf = io.BytesIO()
    arr = np.empty((), dtype=object)
    arr[()] = obj
    np.lib.format.write_array_header_1_0(
        buf, {"descr": "|O", "fortran_order": False, "shape": ()})
    pickle.dump(arr, buf, protocol=protocol)
    with zipfile.ZipFile(path, "w") as z:
        v = io.BytesIO()
        np.lib.format.write_array(v, np.asanyarray(1))
        z.writestr("version.npy", v.getvalue())
        z.writestr("arrays.npy", buf.getvalue())
        sd = io.BytesIO()
        np.lib.format.write_array(sd, np.array({"dt": 1.0, "t": 0.0, "count": 0}),
                                  allow_pickle=True)
        z.writestr("solver_data.npy", sd.getvalue())


def test_reader_refuses_other_globals(tmp_path):
    """Only numpy array/dtype/scalar constructors resolve; any other global in
    the pickle stops the read before anything is built from it."""
    bad = tmp_path / "bad_0.npz"
    _npz_with_arrays_object(bad, collections.OrderedDict(fluid={"x": np.zeros(2)}))
    with pytest.raises(pickle.UnpicklingError, match="collections.OrderedDict"):
        utils.read_output(bad)


def test_reader_accepts_numpy_1_module_names_and_protocol_2(tmp_path):
    """Files written by numpy 1.x name numpy.core.multiarray; protocol 2 (the
    highest Python 2 writes) spells globals as text."""
    p = tmp_path / "old_0.npz"
    data = {"fluid": {"x": np.arange(3.0), "tag": np.zeros(3, np.int64), "c": np.float64(2.0)}}
    _npz_with_arrays_object(p, data, protocol=2)
    with zipfile.ZipFile(p) as z:
        members = {n: z.read(n) for n in z.namelist()}
    raw = members["arrays.npy"]
    if b"numpy._core.multiarray" in raw:
        raw = raw.replace(b"numpy._core.multiarray", b"numpy.core.multiarray")
    assert b"numpy.core.multiarray" in raw
    members["arrays.npy"] = raw
    with zipfile.ZipFile(p, "w") as z:
        for n, b in members.items():
            z.writestr(n, b)
    with warnings.catch_warnings():
        # importing numpy.core under numpy 2 warns that the name is deprecated
        warnings.simplefilter("ignore", DeprecationWarning)
        got = utils.read_output(p)
    assert np.array_equal(got["arrays"]["fluid"]["x"], np.arange(3.0))
    assert got["arrays"]["fluid"]["c"] == 2.0


def test_wrong_version_and_missing_version(tmp_path):
    p = tmp_path / "v2_0.npz"
    np.savez(p, version=2, arrays={}, solver_data={})
    with pytest.raises(RuntimeError, match="Version not understood"):
        utils.read_output(p)
    q = tmp_path / "nov_0.npz"
    np.savez(q, arrays={})
    with pytest.raises(RuntimeError, match="No version number"):
        utils.read_output(q)


def test_writer_reports_errors(tmp_path):
    fl, wl = _arrays()
    (tmp_path / "blocker").write_text("")
    s = _solver_with((fl, wl), tmp_path / "blocker")
    with pytest.raises(OSError, match="already exists"):
        s.dump_output(1e-5)
    w = utils.OutputWriter()
    w.submit(str(tmp_path / "no_such_dir" / "x_0.npz"), {"fluid": {"x": fl.get_carray("x")}},
             {"dt": 0.0, "t": 0.0, "count": 0})
    with pytest.raises(RuntimeError, match="writing an output file failed"):
        w.wait()
