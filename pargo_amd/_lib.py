"""ctypes binding of libpargo.so (the C ABI declared in include/pargo.h).

The library is loaded from pargo_amd/_build/libpargo.so.  There is no
fallback: if the library is missing, or if a GPU entry point is called on a
machine without a HIP device, the call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpargo.so")

PARGO_OK = 0
PARGO_E_INVALID = -1
PARGO_E_DEVICE = -2
PARGO_E_NOMEM = -3
PARGO_E_UNSUPPORTED = -4
PARGO_E_NONTERM = -5

SLOPE_METHODS = {"FD": 0, "FFD": 1, "MD": 2, "SD": 3, "SFD": 4, "TFD": 5, "TFDW": 6}
UNIT_TAN = 0
UNIT_DEGREES = 1
D8_FLAT_256 = 0
D8_FLAT_8 = 255
FLOW_EXCL = 0x10
FLOW_INERT = 253
FLOW_NODATA = 254
FLOW_NONE = 255
SCA_REFERENCE = 0
SCA_CONVERGE = 1

# Every exported symbol of include/pargo.h with (restype, argtypes).
_c = ctypes
_P = _c.c_void_p
_SIGS = {
    "pargo_abi_version": (_c.c_int, []),
    "pargo_create": (_c.c_int, [_c.c_int, _c.POINTER(_P)]),
    "pargo_destroy": (_c.c_int, [_P]),
    "pargo_last_error": (_c.c_char_p, [_P]),
    "pargo_set_stream": (_c.c_int, [_P, _P]),
    "pargo_own_stream": (_P, [_P]),
    "pargo_sync": (_c.c_int, [_P]),
    "pargo_alloc": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_P)]),
    "pargo_free": (_c.c_int, [_P, _P]),
    "pargo_alloc_host": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_P)]),
    "pargo_free_host": (_c.c_int, [_P, _P]),
    "pargo_h2d": (_c.c_int, [_P, _P, _P, _c.c_size_t]),
    "pargo_d2h": (_c.c_int, [_P, _P, _P, _c.c_size_t]),
    "pargo_slope": (_c.c_int, [_P, _P, _P, _c.c_int, _c.c_int, _P, _P]),
    "pargo_d8": (_c.c_int, [_P, _P, _P, _c.c_int, _P, _P]),
    "pargo_slope_d8": (_c.c_int, [_P, _P, _P, _c.c_int, _c.c_int, _c.c_int, _P, _P, _P]),
    "pargo_slope_d8_noalloc": (_c.c_int, [_P, _P, _P, _c.c_int, _c.c_int, _c.c_int, _P, _P, _P]),
    "pargo_idw_grid": (_c.c_int, [_P, _P, _c.c_int, _c.c_float, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int),
                                  _c.POINTER(_c.c_double), _c.POINTER(_c.c_double)]),
    "pargo_idw_extent": (_c.c_int, [_P, _P, _c.c_int, _c.c_float, _P, _c.POINTER(_c.c_int),
                                    _c.POINTER(_c.c_int)]),
    "pargo_idw_prepare": (_c.c_int, [_P, _P, _P, _P, _c.c_int, _c.c_float, _c.c_double, _c.POINTER(_P)]),
    "pargo_idw_plan_free": (_c.c_int, [_P, _P]),
    "pargo_idw_set_values": (_c.c_int, [_P, _P, _P]),
    "pargo_idw_set_values_device": (_c.c_int, [_P, _P, _P]),
    "pargo_idw_run": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_double, _P, _c.c_int, _c.c_int, _c.c_int,
                                 _P, _P]),
    "pargo_idw_order_tiles": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_double, _P, _c.c_int, _c.c_int,
                                         _c.c_int, _c.c_int]),
    "pargo_idw_load": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _P, _c.c_int, _c.c_int, _c.c_double, _c.c_double,
                                  _P, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "pargo_idw_work": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_double, _P, _c.c_int, _c.c_int, _c.c_int,
                                  _P, _P]),
    "pargo_intensity_load": (_c.c_int, [_P, _P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_double,
                                        _c.c_double, _c.c_int, _P, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "pargo_load_row_bands": (_c.c_int, [_P, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _P]),
    "pargo_geotiff_info": (_c.c_int, [_c.c_char_p, _P]),
    "pargo_geotiff_read": (_c.c_int, [_c.c_char_p, _P, _P]),
    "pargo_geotiff_write": (_c.c_int, [_c.c_char_p, _P, _P, _c.c_int, _c.c_char_p]),
    "pargo_points_read": (_c.c_int, [_c.c_char_p, _c.c_int, _c.POINTER(_c.c_int), _P, _P, _P]),
    "pargo_d8_classify": (_c.c_int, [_P, _P, _P, _P]),
    "pargo_sca_d8": (_c.c_int, [_P, _P, _P, _P, _c.c_int, _P]),
    "pargo_pit_remove": (_c.c_int, [_P, _P, _P, _P, _c.POINTER(_c.c_int)]),
    "pargo_slope_adjoint": (_c.c_int, [_P, _P, _P, _P, _P]),
    "pargo_idw_adjoint": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_double, _P, _c.c_int, _c.c_int,
                                     _c.c_int, _P, _P]),
    "pargo_batch_slope_d8": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "pargo_batch_order": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "pargo_batch_autotune": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_c.c_int),
                                        _c.POINTER(_c.c_int)]),
    "pargo_synth_dem": (_c.c_int, [_c.c_uint64, _c.c_int, _c.c_int, _P, _c.c_int]),
    "pargo_synth_plateaus": (_c.c_int, [_c.c_uint64, _c.c_int, _c.c_int, _P, _c.c_double]),
    "pargo_synth_nodata_discs": (_c.c_int, [_c.c_uint64, _c.c_int, _c.c_int, _P, _c.c_int, _c.c_double,
                                            _c.c_float, _c.POINTER(_c.c_int64)]),
    "pargo_synth_samples": (_c.c_int, [_c.c_uint64, _c.c_int, _c.c_double, _c.c_double, _c.c_double,
                                       _c.c_double, _c.c_double, _P, _P, _P]),
}


class PargoGrid(ctypes.Structure):
    """pargo_grid (include/pargo.h)."""
    _fields_ = [("rows", ctypes.c_int32), ("cols", ctypes.c_int32), ("pitch", ctypes.c_int64),
                ("cellsize", ctypes.c_double), ("nodata", ctypes.c_double)]


class RasterInfo(ctypes.Structure):
    """pargo_raster_info (include/pargo.h)."""
    _fields_ = [("rows", ctypes.c_int32), ("cols", ctypes.c_int32), ("transform", ctypes.c_double * 6),
                ("has_nodata", ctypes.c_int32), ("nodata", ctypes.c_double), ("bits", ctypes.c_int32),
                ("sample_format", ctypes.c_int32)]


def check_host(rc, what):
    """Raises PargoError for a failed context-free (host-only) call."""
    if rc != PARGO_OK:
        msg = lib().pargo_last_error(None)
        raise PargoError(rc, "%s: %s" % (what, msg.decode() if msg else ""))
    return rc


class PargoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pargo error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Loads libpargo.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpargo.so not found at %s; run python -m pargo_amd.build" % LIB_PATH)
        # torch ships its own HIP runtime (soname libamdhip64.so.7).  Load it
        # first so that libpargo binds to that same runtime by soname: two
        # HIP runtimes in one process do not share the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue  # reported by missing_symbols()
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def missing_symbols():
    """Entry points of include/pargo.h that the loaded library lacks."""
    L = lib()
    return [n for n in _SIGS if getattr(L, n, None) is None]


def row_pitch(t, name="raster"):
    """Row pitch in elements of a 2-D row-major tensor, refusing layouts the
    kernels cannot address: rows must be contiguous (stride 1 along a row)
    and must not overlap (stride(0) >= cols when there is more than one row;
    a broadcast tensor such as x.expand(rows, cols) has stride(0) == 0).
    Checked before any device work, so a bad layout never reaches a launch."""
    if t.dim() != 2:
        raise ValueError("%s must be 2-D, got %d dimensions" % (name, t.dim()))
    rows, cols = int(t.shape[0]), int(t.shape[1])
    if cols > 1 and t.stride(1) != 1:
        raise ValueError("%s rows must be contiguous (stride(1) == 1), got strides %s" % (name, tuple(t.stride())))
    if rows > 1 and t.stride(0) < cols:
        raise ValueError("%s rows overlap: stride(0) = %d < %d columns (a broadcast or expanded tensor); "
                         "pass a contiguous copy" % (name, t.stride(0), cols))
    return int(t.stride(0)) if rows > 1 else cols


def check_raster(t, dtype, name):
    """A 2-D tensor of `dtype` with addressable rows (row_pitch) on a HIP
    device; the layout is checked first so the error names it on any device."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s must be a %s tensor, got %s" % (name, dtype, t.dtype))
    row_pitch(t, name)
    if t.device.type != "cuda":
        raise TypeError("%s must be on a HIP device" % name)


def _ptr(t):
    """Device or host pointer of a torch tensor / numpy array / None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    if hasattr(t, "ctypes"):
        return ctypes.c_void_p(t.ctypes.data)
    return ctypes.c_void_p(int(t))


class Context:
    """A pargo_ctx bound to one device.  By default work is queued on torch's
    current stream of that device, so results are ordered with torch ops."""

    def __init__(self, device=0, use_torch_stream=True):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.pargo_create(int(device), ctypes.byref(h))
        if rc != PARGO_OK:
            msg = L.pargo_last_error(None)
            raise PargoError(rc, "pargo_create(%d): %s" % (device, msg.decode() if msg else ""))
        self.h = h
        self.device = int(device)
        self.use_torch_stream = use_torch_stream

    def close(self):
        if self.h:
            lib().pargo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != PARGO_OK:
            msg = lib().pargo_last_error(self.h)
            raise PargoError(rc, msg.decode() if msg else "")
        return rc

    def bind_stream(self):
        if self.use_torch_stream:
            import torch
            s = torch.cuda.current_stream(self.device).cuda_stream
            self.check(lib().pargo_set_stream(self.h, ctypes.c_void_p(s)))

    def sync(self):
        self.check(lib().pargo_sync(self.h))


_default_ctx = {}


def default_context(device=0):
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = Context(device)
        _default_ctx[device] = ctx
    return ctx
