"""Output files of the solver: the version-1 ``.npz`` layout and its reader
(solver/utils.py:22-341, solver/solver.py:605-747).

A version-1 file is an uncompressed zip archive of three ``.npy`` members:

  * ``version``     -- the integer 1 (0-d array);
  * ``arrays``      -- a 0-d object array holding ``{array name: {property
    name: 1-d array | constant}}``;
  * ``solver_data`` -- a 0-d object array holding ``{"dt", "t", "count"}``.

The two dictionaries are stored the way ``numpy.savez`` stores any Python
object, as a pickle inside the ``.npy`` member, which is what
solver.py:696-699 writes. Reading them back therefore means unpickling. The
reader here does that with an unpickler that resolves only the numpy
constructors an array, a dtype or a numpy scalar needs and refuses every
other global, so a file cannot make the reader call anything else; dicts,
lists, strings and Python numbers need no globals at all.
"""
import io
import os
import pickle
import zipfile

import numpy as np
from numpy.lib import format as npy_format

FORMAT_VERSION = 1

# (module, name) pairs an ndarray / dtype / numpy scalar pickle refers to,
# under the module names numpy 1.x and 2.x write; protocol-2 pickles written
# by Python 3 also spell byte strings as _codecs.encode(text, "latin1")
_ALLOWED_GLOBALS = {
    ("_codecs", "encode"),
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"),
    ("numpy._core.multiarray", "scalar"),
}


class _ArrayOnlyUnpickler(pickle.Unpickler):
    """Unpickler that builds numpy arrays, dtypes and scalars and nothing else."""

    def find_class(self, module, name):
        if (module, name) in _ALLOWED_GLOBALS:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(
            f"output file refers to {module}.{name}; only numpy arrays, dtypes and "
            f"scalars are accepted in a solver output file")


def check_array(x, y):
    """True when x and y agree to an absolute 1e-16 (utils.py:22-25)."""
    return np.allclose(x, y, atol=1e-16, rtol=0)


def mkdir(newdir):
    """Create a directory and its parents; an existing directory is fine, a
    file in the way is an error (utils.py:263-290)."""
    if os.path.isdir(newdir):
        return
    if os.path.isfile(newdir):
        raise OSError(f"a file with the same name as the desired dir, '{newdir}', "
                      f"already exists.")
    os.makedirs(newdir, exist_ok=True)


def savez(file, *args, **kwds):
    """Uncompressed ``.npz`` with zip64 enabled (utils.py:34-104, :123-170)."""
    np.savez(file, *args, **kwds)


def savez_compressed(file, *args, **kwds):
    """Deflated ``.npz`` (utils.py:106-121)."""
    np.savez_compressed(file, *args, **kwds)


def _read_member(raw):
    """One ``.npy`` member: plain arrays through numpy without pickles, object
    arrays through the array-only unpickler."""
    fp = io.BytesIO(raw)
    major, _minor = npy_format.read_magic(fp)
    if major == 1:
        shape, fortran, dtype = npy_format.read_array_header_1_0(fp)
    elif major in (2, 3):
        shape, fortran, dtype = npy_format.read_array_header_2_0(fp)
    else:
        raise RuntimeError(f"unsupported .npy format version {major}")
    if not dtype.hasobject:
        fp.seek(0)
        return npy_format.read_array(fp, allow_pickle=False)
    # Python 2 writers stored byte strings: latin1 maps them one to one
    return _ArrayOnlyUnpickler(fp, encoding="latin1").load()


def read_output(fname):
    """The contents of a version-1 output file as plain numpy data:
    ``{"version": 1, "arrays": {name: {prop: array | constant}},
    "solver_data": {...}}``."""
    with zipfile.ZipFile(fname) as z:
        members = {os.path.splitext(n)[0]: n for n in z.namelist()}
        if "version" not in members:
            raise RuntimeError("Wrong file type! No version number recorded.")
        version = int(_read_member(z.read(members["version"])))
        if version != FORMAT_VERSION:
            raise RuntimeError("Version not understood!")
        out = {"version": version}
        for key in ("arrays", "solver_data"):
            if key not in members:
                raise RuntimeError(f"version-1 output file without '{key}'")
            obj = _read_member(z.read(members[key]))
            if isinstance(obj, np.ndarray) and obj.dtype.hasobject and obj.shape == ():
                obj = obj[()]
            if not isinstance(obj, dict):
                raise RuntimeError(f"'{key}' of a version-1 output file must hold a dict")
            out[key] = obj
    return out


def load(fname, cl_precision="double", device=None):
    """Load an output file (utils.py:302-341): ``{"arrays": {name:
    ParticleArray}, "solver_data": {...}}``.

    The reference builds the arrays with ``cl_precision="single"``, which only
    matters to its OpenCL path; here ``cl_precision`` selects the device
    arithmetic, so it is a parameter with the FP64 default. Property types
    follow the stored dtypes (float64 -> double, int64 -> long, int32 ->
    int, float32 -> float); the particle type is not stored in the file and
    comes back as Fluid, as it does in the reference."""
    from ..base.particle_array import ParticleArray
    data = read_output(fname)
    ret = {"arrays": {}, "solver_data": data["solver_data"]}
    for name, props in data["arrays"].items():
        pa = ParticleArray(name=name, cl_precision=cl_precision, device=device)
        pa.initialize(**property_infos(props))
        pa.constants.update(constants_of(props))
        ret["arrays"][name] = pa
    return ret


_NP_TYPE_NAMES = {np.dtype(np.float64): "double", np.dtype(np.int64): "long",
                  np.dtype(np.int32): "int", np.dtype(np.float32): "float"}


def _is_property(value):
    return isinstance(value, np.ndarray) and value.ndim == 1


def property_infos(props):
    """``add_property`` descriptions of the per-particle entries of one stored
    array (1-d arrays; everything else is a constant)."""
    infos = {}
    for prop, value in props.items():
        if not _is_property(value):
            continue
        type_name = _NP_TYPE_NAMES.get(value.dtype)
        if type_name is None:
            raise RuntimeError(f"property {prop} has unsupported dtype {value.dtype}")
        infos[prop] = {"data": value, "type": type_name}
    return infos


def constants_of(props):
    return {k: v for k, v in props.items() if not _is_property(v)}


def output_counts(directory, fname):
    """Iteration counts of the files ``<fname>_<count>.npz`` in ``directory``,
    ascending (solver.py:720, :723)."""
    if not os.path.isdir(directory):
        return []
    counts = set()
    prefix = fname + "_"
    for entry in os.listdir(directory):
        if entry.startswith(prefix) and entry.endswith(".npz"):
            tail = entry.rsplit("_", 1)[1][:-4]
            if tail.isdigit():
                counts.add(tail)
    return sorted(counts, key=int)


class OutputWriter:
    """Writes output files without holding up the time loop.

    ``submit`` queues device-to-host copies of the dumped properties into
    pinned host buffers on the current stream, which is the stream the
    operators run on. The next step's kernels therefore cannot overwrite a
    property before its copy has been taken. A worker thread waits for the
    copies and writes the file. At most ``max_pending`` files are in flight;
    a further ``submit`` first waits for the oldest one. ``wait`` returns when
    every file is on disk and raises the first error a worker met."""

    def __init__(self, max_pending=2):
        import threading
        self._threading = threading
        self.max_pending = max(1, int(max_pending))
        self._pending = []
        self._errors = []

    def _host_copy(self, t):
        import torch
        if t.is_cuda:
            host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
            host.copy_(t, non_blocking=True)
            return host
        return t.detach().clone()

    def submit(self, path, arrays, solver_data, version=FORMAT_VERSION):
        """``arrays``: {name: {prop: tensor | constant}}."""
        import torch
        while len(self._pending) >= self.max_pending:
            self._pending.pop(0).join()
        self._raise()
        host, on_device = {}, False
        for name, props in arrays.items():
            host[name] = {}
            for prop, value in props.items():
                if torch.is_tensor(value):
                    on_device = on_device or value.is_cuda
                    host[name][prop] = self._host_copy(value)
                else:
                    host[name][prop] = value
        event = None
        if on_device:
            event = torch.cuda.Event()
            event.record()
        solver_data = dict(solver_data)

        def write():
            try:
                if event is not None:
                    event.synchronize()
                out = {name: {p: (v.numpy() if torch.is_tensor(v) else v)
                              for p, v in props.items()} for name, props in host.items()}
                savez(path, version=version, arrays=out, solver_data=solver_data)
            except BaseException as e:  # handed to the caller by wait()
                self._errors.append(e)

        th = self._threading.Thread(target=write, name="pysph-output", daemon=True)
        th.start()
        self._pending.append(th)

    def wait(self):
        while self._pending:
            self._pending.pop(0).join()
        self._raise()

    def _raise(self):
        if self._errors:
            err = self._errors.pop(0)
            self._errors.clear()
            raise RuntimeError(f"writing an output file failed: {err}") from err
