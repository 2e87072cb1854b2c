"""ctypes binding of libsphhip.so (the C ABI declared in include/sphhip.h).

The library is the only compute path of pysph_amd: there is no CPU
fallback.  Loading fails loudly when the shared object is missing, and the
context fails loudly when no HIP device is visible.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphhip.so")

SPH_OK = 0
SPH_EINVAL = -1
SPH_ENEEDCAP = -2
SPH_EHIP = -3
SPH_EDOMAIN = -4
SPH_ENOMEM = -5

ORDER_MORTON = 0
ORDER_LINEAR = 1

MAX_PROPS = 16
MAX_STEP_PROPS = MAX_PROPS

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_dbl = ctypes.c_double


class SphGrid(ctypes.Structure):
    _fields_ = [("cell_size", _dbl), ("max_h", _dbl), ("min_h", _dbl),
                ("cmin", ctypes.c_int32 * 3), ("dims", ctypes.c_int32 * 3),
                ("bits", ctypes.c_int32 * 3), ("order", ctypes.c_int32),
                ("ncells", _i64)]


class PropListF32(ctypes.Structure):
    _fields_ = [("nprops", ctypes.c_int32),
                ("dst_stride", _i64),
                ("src", _vp * MAX_PROPS),
                ("dst", _vp * MAX_PROPS)]


class StepProps(ctypes.Structure):
    _fields_ = [("nprops", ctypes.c_int32),
                ("x", _vp * MAX_PROPS),
                ("x0", _vp * MAX_PROPS),
                ("a", _vp * MAX_PROPS)]


class PropList(ctypes.Structure):
    _fields_ = [("nprops", ctypes.c_int32),
                ("dst_stride", _i64),
                ("src", _vp * MAX_PROPS),
                ("dst", _vp * MAX_PROPS)]


MAX_SRC = 4


class Sources(ctypes.Structure):
    """sph_sources (include/sphhip.h)."""
    _fields_ = [("nsrc", ctypes.c_int32), ("counts", _vp * MAX_SRC), ("slice_off", _vp * MAX_SRC),
                ("indices", _vp * MAX_SRC), ("s_rec", _vp * MAX_SRC), ("ns", _i64 * MAX_SRC),
                ("s_pm", _vp * MAX_SRC), ("s_cell_start", _vp * MAX_SRC),
                ("same", ctypes.c_int32 * MAX_SRC)]


class NnpsDevArgs(ctypes.Structure):
    """sph_nnps_dev_args (include/sphhip.h): one device-side list build."""
    _fields_ = [("f32", ctypes.c_int32), ("same_array", ctypes.c_int32), ("d_rec", _vp),
                ("d_x", _vp), ("d_y", _vp), ("d_z", _vp), ("d_perm", _vp),
                ("d_cell_start", _vp), ("d_cells", _vp), ("nd_cells_dev", _vp),
                ("max_cells", _i64), ("nd_total", _i64), ("nd_real", _i64),
                ("s_cell_start", _vp), ("s_rec", _vp), ("ns", _i64),
                ("radius_scale", ctypes.c_double), ("counts", _vp), ("slice_off", _vp),
                ("indices", _vp), ("cap", _i64), ("totals", _vp), ("status", _vp),
                ("s_near", _vp), ("near_reach", ctypes.c_int32)]


PAIR_PRESSURE_GRADIENT = 1
PAIR_ARTIFICIAL_VISCOSITY = 2
PAIR_MOMENTUM = 3
PAIR_BOUNDARY_FORCE = 4


class PairParams(ctypes.Structure):
    """sph_pair_params (include/sphhip.h)."""
    _fields_ = [("op", ctypes.c_int32), ("with_correction", ctypes.c_int32)] + [
        (k, ctypes.c_double) for k in ("alpha", "beta", "eta", "deltap", "n", "eps", "delp")]


class WallFrames(ctypes.Structure):
    """sph_wall_frames (include/sphhip.h)."""
    _fields_ = [("s_perm", _vp * MAX_SRC)] + [
        (k, _vp * MAX_SRC) for k in ("tx", "ty", "tz", "nx", "ny", "nz")]


class BatchDesc(ctypes.Structure):
    """sph_batch_desc (include/sphhip.h)."""
    _fields_ = ([("nsims", ctypes.c_int32), ("dim", ctypes.c_int32), ("n", _i64),
                 ("sim_off", _vp), ("sim_of", _vp)]
                + [(k, _vp) for k in ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs",
                                      "x0", "y0", "z0", "u0", "v0", "w0",
                                      "ax", "ay", "az", "au", "av", "aw", "dt_fac",
                                      "dt_sim")]
                + [(k, _dbl) for k in ("co", "ro", "gamma", "alpha", "beta", "eta",
                                       "radius_scale")]
                + [("key_cap", _i64), ("max_cells", _i64), ("list_cap", _i64),
                   ("maxch", ctypes.c_int32), ("f32", ctypes.c_int32),
                   ("uniform_h", ctypes.c_int32), ("op_threads", ctypes.c_int32),
                   ("op_blocks_per_cu", ctypes.c_int32), ("nbr_tile", ctypes.c_int32),
                   ("cells_per_block", ctypes.c_int32), ("work", _vp),
                   ("work_bytes", _i64),
                   ("status", _vp), ("pairs", _vp), ("persistent", ctypes.c_int32),
                   ("staged", ctypes.c_int32)])


# name -> (restype, argtypes)
_SIGNATURES = {
    "sph_version": (ctypes.c_char_p, []),
    "sph_ctx_create": (_i32, [ctypes.POINTER(_vp), _i32, _vp]),
    "sph_ctx_destroy": (_i32, [_vp]),
    "sph_ctx_set_stream": (_i32, [_vp, _vp]),
    "sph_ctx_synchronize": (_i32, [_vp]),
    "sph_last_error": (ctypes.c_char_p, [_vp]),
    "sph_grid_build": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _i32, _i64,
                              _i32, ctypes.POINTER(SphGrid)]),
    "sph_grid_check": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, ctypes.POINTER(SphGrid), _i32,
                              _vp]),
    "sph_bin_array": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                             _vp, _vp, ctypes.POINTER(_i64)]),
    "sph_gather_sorted": (_i32, [_vp, _i64, _vp, ctypes.POINTER(PropList)]),
    "sph_scatter_sorted": (_i32, [_vp, _i64, _vp, ctypes.POINTER(PropList)]),
    "sph_copy_many": (_i32, [_vp, _i64, ctypes.POINTER(PropList)]),
    "sph_fill_many": (_i32, [_vp, _i64, _dbl, ctypes.POINTER(PropList)]),
    "sph_scatter_sorted2": (_i32, [_vp, _i64, _vp, ctypes.POINTER(PropList),
                                   ctypes.POINTER(PropList)]),
    "sph_scatter_sorted_f32": (_i32, [_vp, _i64, _vp, ctypes.POINTER(PropListF32)]),
    "sph_nnps_build": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _vp, _i64, _i64, _i64,
                              _vp, _vp, _i64, _dbl, _i32, _vp, _vp, _vp, _i64,
                              ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "sph_nnps_export": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sph_nnps_build_varh": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _vp, _vp, _i64,
                                   _i64, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _dbl, _dbl, _vp,
                                   _vp, _vp, _i64, ctypes.POINTER(_i64),
                                   ctypes.POINTER(_i64)]),
    "sph_nnps_build_varh_f32": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64,
                                       _i64, _dbl, _dbl, _vp, _vp, _vp, _i64,
                                       ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "sph_bin_array_dev": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _i64, _vp, _vp,
                                 _vp, _vp, _vp, _vp]),
    "sph_nnps_build_dev": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _vp, _vp, _i64,
                                  _i64, _i64, _vp, _vp, _i64, _dbl, _i32, _vp, _vp, _vp, _i64,
                                  _vp, _vp, _vp, _i32]),
    "sph_nnps_build_dev_pair": (_i32, [_vp, ctypes.POINTER(SphGrid),
                                       ctypes.POINTER(NnpsDevArgs),
                                       ctypes.POINTER(NnpsDevArgs)]),
    "sph_cell_near": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                             _vp, _i32, _vp]),
    "sph_nnps_build_f32_dev": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _dbl,
                                      _i32, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i32]),
    "sph_summation_density": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32,
                                     _i32, _vp, _vp, _i64]),
    "sph_tait": (_i32, [_vp, _i64, _i64, _vp, _vp, _i64, _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp,
                        _vp, _vp]),
    "sph_momentum_av": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32,
                               _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp, _vp]),
    "sph_cubic_spline": (_i32, [_vp, _i32, _i64] + [_vp] * 11),
    "sph_summation_density_multi": (_i32, [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(Sources),
                                           _vp, _i32, _i32, _vp, _vp, _i64]),
    "sph_summation_density_multi_f32": (_i32, [_vp, _i32, _i64, _i64, _vp,
                                               ctypes.POINTER(Sources), _vp, _i32, _i32, _vp,
                                               _vp, _i64]),
    "sph_summation_density_tait": (_i32, [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(Sources),
                                          _vp, _i32, _i32, _vp, _vp, _i64, _dbl, _dbl, _dbl,
                                          _i32, _vp, _vp, _vp, _vp, _vp]),
    "sph_summation_density_tait_f32": (_i32, [_vp, _i32, _i64, _i64, _vp,
                                              ctypes.POINTER(Sources), _vp, _i32, _i32, _vp,
                                              _vp, _i64, _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp,
                                              _vp, _vp]),
    "sph_momentum_av_multi": (_i32, [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(Sources), _vp,
                                     _i32, _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp, _vp]),
    "sph_momentum_av_multi_f32": (_i32, [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(Sources),
                                         _vp, _i32, _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp,
                                         _vp]),
    "sph_summation_density_staged": (_i32, [_vp, ctypes.POINTER(SphGrid), _dbl, _i32, _i64, _i64,
                                            _vp, ctypes.POINTER(Sources), _vp, _i32, _i32, _vp,
                                            _vp, _i64]),
    "sph_momentum_av_staged": (_i32, [_vp, ctypes.POINTER(SphGrid), _dbl, _i32, _i64, _i64, _vp,
                                      ctypes.POINTER(Sources), _vp, _i32, _dbl, _dbl, _dbl, _i32,
                                      _vp, _vp, _vp, _vp]),
    "sph_density_rate": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32,
                                _vp]),
    "sph_density_rate_f32": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32,
                                    _i32, _vp]),
    "sph_pair_accel": (_i32, [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(Sources), _vp, _i32,
                              ctypes.POINTER(PairParams), ctypes.POINTER(WallFrames), _i32,
                              _vp, _vp, _vp, _vp]),
    "sph_pair_accel_f32": (_i32, [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(Sources), _vp,
                                  _i32, ctypes.POINTER(PairParams), ctypes.POINTER(WallFrames),
                                  _i32, _vp, _vp, _vp, _vp]),
    "sph_xsph": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _dbl, _i32,
                        _vp, _vp, _vp]),
    "sph_xsph_f32": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _dbl,
                            _i32, _vp, _vp, _vp]),
    "sph_extrema": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, ctypes.POINTER(_dbl)]),
    "sph_extrema_noalloc": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _vp]),
    "sph_gravity": (_i32, [_vp, _i64, _i64, _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp]),
    "sph_position_stepping": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sph_stage_step": (_i32, [_vp, _i64, ctypes.POINTER(StepProps), _dbl]),
    "sph_pc_final": (_i32, [_vp, _i64, ctypes.POINTER(StepProps)]),
    "sph_copy": (_i32, [_vp, _i64, _vp, _vp]),
    "sph_gather_sorted_f32": (_i32, [_vp, _i64, _vp, ctypes.POINTER(PropListF32)]),
    "sph_nnps_build_f32": (_i32, [_vp, ctypes.POINTER(SphGrid), _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp, _i64, _i64, _i64, _vp, _vp, _i64, _dbl, _i32, _vp, _vp,
                                  _vp, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "sph_summation_density_f32": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _i32, _i32, _vp, _vp, _i64]),
    "sph_tait_f32": (_i32, [_vp, _i64, _i64, _vp, _vp, _i64, _dbl, _dbl, _dbl, _i32, _vp, _vp,
                            _vp, _vp, _vp]),
    "sph_momentum_av_f32": (_i32, [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                   _i32, _dbl, _dbl, _dbl, _i32, _vp, _vp, _vp, _vp]),
    "sph_batch_workspace_size": (_i32, [_vp, _vp, ctypes.POINTER(_i64)]),
    "sph_batch_rk2_step": (_i32, [_vp, _vp, _dbl]),
    "sph_batch_init_workspace": (_i32, [_vp, _vp]),
    "sph_batch_graph_create": (_i32, [_vp, _vp, _dbl, _i32, ctypes.POINTER(_vp)]),
    "sph_batch_graph_launch": (_i32, [_vp, _vp]),
    "sph_batch_graph_destroy": (_i32, [_vp, _vp]),
    "sph_batch_save_initial": (_i32, [_vp, _vp]),
    "sph_batch_eval": (_i32, [_vp, _vp]),
    "sph_batch_stage": (_i32, [_vp, _vp, _dbl]),
    "sph_batch_vjp": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sph_fill": (_i32, [_vp, _i64, _dbl, _vp]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None
_lib_lock = threading.Lock()


class SphError(RuntimeError):
    """A libsphhip call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"libsphhip error {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    """Load libsphhip.so and declare its signatures (no GPU needed)."""
    global _lib
    with _lib_lock:
        if _lib is not None and path == LIB_PATH:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it with "
                f"`python -c 'import __graft_entry__ as g; g.build()'` "
                f"(pysph_amd has no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path == LIB_PATH:
            _lib = lib
        return lib


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


class Context:
    """One sph_ctx per device, bound to the caller's current torch stream, or
    -- when `stream` is given -- to that torch stream for good (a context of
    its own, e.g. for graph capture)."""

    def __init__(self, device=None, stream=None):
        if not torch.cuda.is_available():
            raise RuntimeError("pysph_amd needs a HIP device (torch.cuda.is_available() is "
                               "False); it has no CPU fallback")
        self.lib = load_library()
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        idx = self.device.index if self.device.index is not None else 0
        self.device_index = idx
        self.fixed_stream = stream is not None
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        h = _vp()
        rc = self.lib.sph_ctx_create(ctypes.byref(h), idx, _vp(self.stream.cuda_stream))
        if rc != SPH_OK:
            raise SphError(rc, f"sph_ctx_create(device={idx}) failed")
        self.handle = h

    def sync_stream(self):
        """Follow torch's current stream (torch ops and our kernels stay ordered)."""
        if self.fixed_stream:
            return
        s = torch.cuda.current_stream(self.device)
        if s.cuda_stream != self.stream.cuda_stream:
            self.stream = s
            self.check(self.lib.sph_ctx_set_stream(self.handle, _vp(s.cuda_stream)))

    def check(self, rc, allow=()):
        if rc == SPH_OK or rc in allow:
            return rc
        msg = self.lib.sph_last_error(self.handle)
        raise SphError(rc, msg.decode() if msg else "")

    def synchronize(self):
        self.check(self.lib.sph_ctx_synchronize(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sph_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def context(device=None):
    """The shared Context for `device` (default cuda:0)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    key = dev.index if dev.index is not None else 0
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = Context(torch.device("cuda", key))
        _contexts[key] = ctx
    ctx.sync_stream()
    return ctx
