"""ctypes binding of libtgrav.so (the C ABI declared in include/tgrav.h).

The library must have been built (``torch_amd.build.build()``); importing this module
never falls back to anything else.  A missing or stale library is an error.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtgrav.so")

PLAIN_FP64, MIXED = 0, 1
HOST, DEVICE = 0, 1

OK, EINVAL, ENOSOURCES, EHIP, ENONFINITE, ENOMEM, ELIMIT = 0, -1, -2, -3, -4, -5, -6
ERROR_NAMES = {EINVAL: "EINVAL", ENOSOURCES: "ENOSOURCES", EHIP: "EHIP",
               ENONFINITE: "ENONFINITE", ENOMEM: "ENOMEM", ELIMIT: "ELIMIT"}

# (name, restype, argtypes); P = double pointer passed as c_void_p (host or device address)
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int
_D = ctypes.c_double


class Problem(ctypes.Structure):
    """tgrav_problem (one problem of a batched evaluation; device pointers)."""
    _fields_ = [("n_sources", _I64), ("sx", _P), ("sy", _P), ("sz", _P), ("sm", _P), ("se", _P),
                ("n_targets", _I64), ("tx", _P), ("ty", _P), ("tz", _P), ("te", _P),
                ("ax", _P), ("ay", _P), ("az", _P), ("phi", _P),
                ("vx", _P), ("vy", _P), ("vz", _P),
                ("n_target_blocks", _I64), ("tbx", _P), ("tby", _P), ("tbz", _P), ("tbd", _P),
                ("tbe", _P),
                ("n_source_blocks", _I64), ("sbx", _P), ("sby", _P), ("sbz", _P), ("sbd", _P),
                ("sbe", _P)]


SIGNATURES = [
    ("tgrav_version", _I32, []),
    ("tgrav_last_error", ctypes.c_char_p, []),
    ("tgrav_create", _I32, [ctypes.POINTER(_P), _I32, ctypes.c_char_p]),
    ("tgrav_tune_cache_entries", _I32, [ctypes.c_char_p]),
    ("tgrav_destroy", _I32, [_P]),
    ("tgrav_set_stream", _I32, [_P, _P]),
    ("tgrav_use_own_stream", _I32, [_P]),
    ("tgrav_get_stream", _P, [_P]),
    ("tgrav_synchronize", _I32, [_P]),
    ("tgrav_set_constant", _I32, [_P, _D]),
    ("tgrav_set_validation", _I32, [_P, _I32]),
    ("tgrav_set_config", _I32, [_P, _I32, _I32, _I32]),
    ("tgrav_num_variants", _I32, [_I32]),
    ("tgrav_variant_info", _I32, [_I32, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                                  ctypes.POINTER(_I32)]),
    ("tgrav_set_sources", _I32, [_P, _I64, _P, _P, _P, _P, _P, _I32]),
    ("tgrav_set_source_blocks", _I32, [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _I32]),
    ("tgrav_eval", _I32, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32]),
    ("tgrav_eval_kick", _I32, [_P, _I64, _P, _P, _P, _P, _D, _P, _P, _P, _P, _P, _P, _P,
                               _I32, _I32]),
    ("tgrav_kick", _I32, [_P, _I64, _D, _P, _P, _P, _P, _P, _P, _I32]),
    ("tgrav_eval_static", _I32, [_P, _I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P,
                                 _I32]),
    ("tgrav_eval_blocks", _I32, [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D,
                                 _P, _P, _P, _I32, _I32]),
    ("tgrav_bidirectional", _I32, [_P, _I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P,
                                   _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32]),
    ("tgrav_bidirectional_blocks", _I32, [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P,
                                          _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32]),
    ("tgrav_set_block_kernel", _I32, [_P, _I32]),
    ("tgrav_set_column_waves", _I32, [_P, _I32]),
    ("tgrav_get_column_waves", _I32, [_P]),
    ("tgrav_set_softening", _I32, [_P, _I32, _D, _I32]),
    ("tgrav_set_periodic", _I32, [_P, _I32, _D, _D, _D, _I32]),
    ("tgrav_pm_deposit", _I32, [_P, _I64, _P, _P, _P, _P, _P, _I64, _I32, _I32, _P, _P, _P, _P, _P,
                                _I32, _I32]),
    ("tgrav_pm_gradient", _I32, [_P, _I64, _I32, _I32, _P, _P, _P, _P, _P, _D, _P, _P, _P, _I32]),
    ("tgrav_pm_interpolate", _I32, [_P, _I64, _P, _P, _P, _P, _I64, _I32, _I32, _P, _P, _P, _P, _P,
                                    _P, _P, _P, _P, _P, _I32]),
    ("tgrav_set_bidir_config", _I32, [_P, _I32, _I32, _I32]),
    ("tgrav_bidir_variant_info", _I32, [_I32, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    ("tgrav_batch_create", _I32, [_P, ctypes.c_int32, ctypes.POINTER(Problem), _I32, _D,
                                  ctypes.POINTER(_P)]),
    ("tgrav_batch_run", _I32, [_P, _I32]),
    ("tgrav_batch_destroy", _I32, [_P]),
    ("tgrav_batch_info", _I32, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_D)]),
    ("tgrav_set_batch_config", _I32, [_P, _D, _I32]),
    ("tgrav_set_batch_scales", _I32, [_P, _D, _D]),
    ("tgrav_batch_partials", _I32, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("tgrav_set_timing", _I32, [_P, _I32]),
    ("tgrav_timing_read", _I32, [_P, _I32, ctypes.POINTER(_D), ctypes.POINTER(_I64),
                                 ctypes.POINTER(_I64)]),
]

_lib = None


class TgravError(RuntimeError):
    """A non-zero return code of the C ABI (AMUSE's ERROR_CODE convention)."""

    def __init__(self, code, where, msg):
        self.code = code
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)}: {msg}")


def _share_torch_hip_runtime():
    """Make libtgrav.so bind to the HIP runtime PyTorch uses, so one process holds one HIP
    runtime whatever the import order.  PyTorch-ROCm ships its own libamdhip64.so with the
    same SONAME as /opt/rocm's; loading torch's copy first (by absolute path, without
    importing torch) makes the dynamic loader reuse it for libtgrav.so, and later makes
    torch find the same file already loaded.  Two runtimes in one process would each open
    the GPU and could not share streams."""
    if os.environ.get("TGRAV_HIP_RUNTIME", "") == "system":
        return None
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return None
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if not os.path.exists(path):
        return None
    ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return path


HIP_RUNTIME = None


def load():
    """Load libtgrav.so (built in-tree).  Raises if it is missing."""
    global _lib, HIP_RUNTIME
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m torch_amd.build`")
    HIP_RUNTIME = _share_torch_hip_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc, where):
    if rc != 0:
        msg = load().tgrav_last_error().decode(errors="replace")
        raise TgravError(rc, where, msg)
    return rc
