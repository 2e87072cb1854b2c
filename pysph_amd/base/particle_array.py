"""ParticleArray: the SoA particle store, one device tensor per property.

Mirrors ``pysph.base.particle_array`` (base/particle_array.pyx:96-1488):
the same constructor, property model (``tag``/``group``/``local``/``pid``
always present; ``get_particle_array`` defaults at :1431-1434), the dirty
flag set when x/y/z are written (:753-754), ``copy_over_properties``
(:1180-1209), ``set_to_zero`` (:1211-1223), ``remove_particles`` with the
reference's swap-with-last removal (:380-435, carray ``remove``), and the
LocalReal-first alignment (:995-1081).

Differences that follow from device storage:
  * property data are torch tensors on one device (float64 for 'double',
    int64 for 'long', int32 for 'int', float32 for 'float');
  * ``get``/attribute access return host numpy COPIES (the reference returns
    views); write with ``set`` (or operate on ``get_carray`` tensors);
  * ``get_carray(name)`` returns the device tensor itself (the analogue of
    the reference's carray).
"""
import numpy as np
import torch

from .. import default_device
from .particle_types import LocalReal, ParticleType

_DTYPES = {
    None: torch.float64,
    "double": torch.float64,
    "long": torch.int64,
    "int": torch.int32,
    "float": torch.float32,
}

_TYPE_NAMES = {torch.float64: "double", torch.int64: "long", torch.int32: "int",
               torch.float32: "float"}


class ParticleArray:
    """Collection of particles stored as one device array per property."""

    def __init__(self, name="", default_particle_tag=LocalReal, particle_type=ParticleType.Fluid,
                 cl_precision="double", constants=None, device=None, **props):
        object.__setattr__(self, "_initialising", True)
        self.device = torch.device(device) if device is not None else default_device()
        self.properties = {}
        self.default_values = {"tag": default_particle_tag, "group": 0, "local": 1, "pid": 0}
        for pname, ptype in (("tag", "long"), ("group", "long"), ("local", "int"),
                             ("pid", "int")):
            self.properties[pname] = torch.zeros(0, dtype=_DTYPES[ptype], device=self.device)
        self.temporary_arrays = {}
        self.constants = dict(constants or {})
        self.time = 0.0
        self.name = name
        self.is_dirty = True
        self.indices_invalid = True
        self.particle_type = particle_type
        self.cl_precision = cl_precision
        self.num_real_particles = 0
        # per-property write counters; device shadows compare against them
        self._version = {}
        object.__setattr__(self, "_initialising", False)
        if props:
            self.initialize(**props)

    # ------------------------------------------------------------------
    # attribute convenience (particle_array.pyx:179-204)
    # ------------------------------------------------------------------
    def __getattr__(self, name):
        d = self.__dict__
        if "properties" in d and name in d["properties"]:
            return self.get(name)
        if "temporary_arrays" in d and name in d["temporary_arrays"]:
            return self.get(name)
        if "constants" in d and name in d["constants"]:
            return d["constants"][name]
        raise AttributeError(f"ParticleArray {d.get('name', '')} has no property {name}")

    def __setattr__(self, name, value):
        d = self.__dict__
        if not d.get("_initialising", True) and "properties" in d and (
                name in d["properties"] or name in d["temporary_arrays"]):
            self.set(**{name: value})
        elif not d.get("_initialising", True) and "constants" in d and name in d["constants"]:
            d["constants"][name] = value
        else:
            object.__setattr__(self, name, value)

    # ------------------------------------------------------------------
    # public interface
    # ------------------------------------------------------------------
    def set_dirty(self, value):
        self.is_dirty = bool(value)

    def touch(self, *props):
        """Record that the named properties were written (all when none are
        given).  Code that writes a tensor from ``get_carray`` directly must
        call this so sorted device copies are refreshed."""
        names = props if props else tuple(self.properties) + tuple(self.temporary_arrays)
        for p in names:
            self._version[p] = self._version.get(p, 0) + 1

    def version(self, prop):
        return self._version.get(prop, 0)

    def set_indices_invalid(self, value):
        self.indices_invalid = bool(value)

    def has_array(self, arr_name):
        return arr_name in self.properties

    def clear(self):
        tag_default = self.default_values["tag"]
        self.properties = {}
        self.default_values = {"tag": tag_default, "group": 0, "local": 1, "pid": 0}
        for pname, ptype in (("tag", "long"), ("group", "long"), ("local", "int"),
                             ("pid", "int")):
            self.properties[pname] = torch.zeros(0, dtype=_DTYPES[ptype], device=self.device)
        self.temporary_arrays = {}
        self.touch()
        self.is_dirty = True
        self.indices_invalid = True

    def set_time(self, time):
        self.time = float(time)

    def get_time(self):
        return self.time

    def set_name(self, name):
        self.name = name

    def set_particle_type(self, particle_type):
        self.particle_type = particle_type

    def initialize(self, **props):
        """particle_array.pyx:288-352"""
        self.clear()
        if not props:
            return
        for prop, info in props.items():
            info = dict(info)
            info["name"] = prop
            self.add_property(info)
        self.align_particles()

    def get_number_of_particles(self):
        for arr in self.properties.values():
            return int(arr.numel())
        return 0

    def get_carray(self, prop):
        """The device tensor of a property (the reference returns its carray)."""
        if prop in self.properties:
            return self.properties[prop]
        return self.temporary_arrays.get(prop)

    def _check_property(self, prop):
        if prop in self.temporary_arrays or prop in self.properties or prop in self.constants:
            return
        raise AttributeError(f"property {prop} not present")

    def get(self, *args, only_real_particles=True):
        """Host numpy copies of the named properties (constants as stored)."""
        if len(args) == 0:
            return None
        result = []
        for arg in args:
            self._check_property(arg)
            arr = self.properties.get(arg)
            if arr is None:
                arr = self.temporary_arrays.get(arg)
            if arr is None:
                result.append(self.constants[arg])
                continue
            if only_real_particles:
                arr = arr[:self.num_real_particles]
            result.append(arr.detach().cpu().numpy().copy())
        return result[0] if len(args) == 1 else tuple(result)

    def set(self, **props):
        """Write property values; x/y/z mark the array dirty (:711-754)."""
        for prop in props:
            self._check_property(prop)
        for prop, value in props.items():
            target = self.properties.get(prop)
            if target is None:
                target = self.temporary_arrays.get(prop)
            if target is None:
                continue
            src = torch.as_tensor(np.asarray(value) if not torch.is_tensor(value) else value)
            if src.dim() == 0:
                target.fill_(src.item())
            else:
                src = src.to(device=target.device, dtype=target.dtype)
                if src.numel() != target.numel():
                    raise ValueError(
                        f"size mismatch setting {prop}: {src.numel()} != {target.numel()}")
                target.copy_(src)
            self.touch(prop)
            if prop == "tag":
                self.align_particles()
            if prop in ("x", "y", "z"):
                self.set_dirty(True)

    def add_property(self, prop_info):
        """particle_array.pyx:765-928 (name, type, default, data)."""
        prop_name = prop_info.get("name")
        data_type = prop_info.get("type")
        data = prop_info.get("data")
        default = prop_info.get("default")
        if prop_name is None:
            raise ValueError("Cannot add property with no name")
        n = self.get_number_of_particles()
        if data is not None and not np.isscalar(data):
            data = np.asarray(data.detach().cpu().numpy() if torch.is_tensor(data) else data)
            ndata = len(data)
        else:
            ndata = 0
            data = None
        if ndata and n and ndata != n:
            raise ValueError("Array sizes incompatible")
        if default is None:
            default = self.default_values.get(prop_name, 0) if prop_name in self.properties \
                else 0
        self.default_values[prop_name] = default
        if data_type is None and data is not None and prop_name not in self.properties:
            if data.dtype in (np.int32, np.int64):
                data_type = "long"
            elif data.dtype == np.float32:
                data_type = "float"
            else:
                data_type = "double"
        if n == 0 and ndata:
            # resize every existing property to the new length with defaults
            for pname in list(self.properties):
                arr = self.properties[pname]
                self.properties[pname] = torch.full((ndata,), self.default_values[pname],
                                                    dtype=arr.dtype, device=self.device)
            n = ndata
            if prop_name == "tag":
                self.num_real_particles = int(np.sum(data == LocalReal))
            else:
                self.num_real_particles = ndata
        if prop_name in self.properties:
            if data is not None:
                arr = self.properties[prop_name]
                arr.copy_(torch.as_tensor(data, dtype=arr.dtype).to(self.device))
                self.touch(prop_name)
            return
        dt = _DTYPES.get(data_type, torch.float64)
        if data is not None:
            t = torch.as_tensor(np.ascontiguousarray(data)).to(device=self.device, dtype=dt)
        else:
            t = torch.full((n,), default, dtype=dt, device=self.device)
        self.properties[prop_name] = t.contiguous()
        self.touch()

    def add_temporary_array(self, arr_name):
        if arr_name in self.temporary_arrays or arr_name in self.properties:
            return
        self.temporary_arrays[arr_name] = torch.zeros(self.get_number_of_particles(),
                                                      dtype=torch.float64, device=self.device)
        self.touch(arr_name)

    def output_property_names(self, all=False):
        """The properties ``get_property_arrays`` returns: every one, or those
        not starting with '_' plus _tmpx/_tmpy/_tmpz (particle_array.pyx:634-650)."""
        names = list(self.properties) if all else [
            p for p in self.properties if not p.startswith("_")] + ["_tmpx", "_tmpy", "_tmpz"]
        return [p for p in names if p in self.properties]

    def get_property_arrays(self, all=False):
        """Host copies of the output properties, then the constants
        (particle_array.pyx:634-650)."""
        ret = {}
        for p in self.output_property_names(all):
            ret[p] = self.properties[p].detach().cpu().numpy().copy()
        ret.update(self.constants)
        return ret

    def align_particles(self):
        """Move LocalReal particles first (particle_array.pyx:995-1081).

        Implemented as a stable partition (real particles keep their order,
        then the others in order)."""
        n = self.get_number_of_particles()
        tag = self.properties["tag"]
        if n == 0:
            self.num_real_particles = 0
            return 0
        real = tag == LocalReal
        nreal = int(real.sum().item())
        self.num_real_particles = nreal
        if nreal == n or bool(real[:nreal].all().item()):
            return 0
        order = torch.cat([torch.nonzero(real).flatten(), torch.nonzero(~real).flatten()])
        for d in (self.properties, self.temporary_arrays):
            for k in list(d):
                d[k] = d[k][order].contiguous()
        self.touch()
        self.is_dirty = True
        self.indices_invalid = True
        return 0

    def remove_particles(self, index_list):
        """Swap-with-last removal of every index (carray remove, particle_array.pyx:380-435)."""
        idx = np.sort(np.asarray(index_list, dtype=np.int64).ravel())
        n = self.get_number_of_particles()
        if len(idx) > n:
            raise ValueError("Number of particles to be removed is greater than number of "
                             "particles in array")
        if len(idx) == 0:
            return
        perm = np.arange(n)
        length = n
        for k in range(len(idx) - 1, -1, -1):
            i = idx[k]
            if i < length:
                perm[i] = perm[length - 1]
                length -= 1
        keep = torch.as_tensor(perm[:length], device=self.device)
        for d in (self.properties, self.temporary_arrays):
            for k in list(d):
                d[k] = d[k][keep].contiguous()
        self.touch()
        self.align_particles()
        self.is_dirty = True
        self.indices_invalid = True

    def copy_over_properties(self, props):
        """dst <- src for every (src: dst) pair (particle_array.pyx:1180-1209)."""
        for prop, dst in props.items():
            self.get_carray(dst).copy_(self.get_carray(prop))
            self.touch(dst)

    def set_to_zero(self, props):
        """particle_array.pyx:1211-1223"""
        for prop in props:
            self.get_carray(prop).zero_()
            self.touch(prop)

    def get_type_name(self, prop):
        return _TYPE_NAMES[self.properties[prop].dtype]

    def __repr__(self):
        return (f"ParticleArray(name={self.name!r}, n={self.get_number_of_particles()}, "
                f"type={self.particle_type}, device={self.device})")


def get_particle_array(cl_precision="double", device=None, **props):
    """Create a particle array with the default properties (particle_array.pyx:1400-1488)."""
    prop_dict = {}
    name = ""
    particle_type = ParticleType.Fluid
    default_props = {"x": 0.0, "y": 0.0, "z": 0.0, "u": 0.0, "v": 0.0, "w": 0.0, "m": 1.0,
                     "h": 1.0, "p": 0.0, "e": 0.0, "rho": 1.0, "cs": 0.0, "_tmpx": 0.0,
                     "_tmpy": 0.0, "_tmpz": 0.0}
    np_ = 0
    constants = {}
    for prop, value in props.items():
        if prop in ("name", "type"):
            continue
        if not isinstance(value, np.ndarray) and not torch.is_tensor(value):
            constants[prop] = value
            continue
        np_ = len(value)
        if prop == "idx":
            prop_dict[prop] = {"data": value, "type": "long"}
        elif prop == "tag":
            prop_dict[prop] = {"data": value, "type": "long"}
        else:
            prop_dict[prop] = {"data": value, "type": "double"}
    for prop, dv in default_props.items():
        if prop not in props:
            prop_dict[prop] = {"name": prop, "type": "double", "default": dv}
    if "idx" not in prop_dict and np_ != 0:
        prop_dict["idx"] = {"name": "idx", "data": np.arange(np_), "type": "long"}
    if "name" in props:
        name = props["name"]
    if "type" in props:
        particle_type = props["type"]
        assert particle_type in (ParticleType.Fluid, ParticleType.Solid, ParticleType.Probe,
                                 ParticleType.Boundary, ParticleType.DummyFluid), \
            "Type not understood!"
    # data-bearing properties first so the defaults size to them
    ordered = dict(sorted(prop_dict.items(), key=lambda kv: "data" not in kv[1]))
    pa = ParticleArray(name=name, particle_type=particle_type, cl_precision=cl_precision,
                       device=device, **ordered)
    pa.constants.update(constants)
    return pa
