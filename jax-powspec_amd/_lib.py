"""ctypes binding of libpowspec_amd.so (include/pk_api.h).

The library is loaded lazily on first use.  If it is missing, every GPU entry point
raises: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

from . import _build

PK_OK = 0
PK_E_INVALID = -1
PK_E_OOM = -2
PK_E_LAUNCH = -3
PK_E_UNSUPPORTED = -4
PK_E_WORKSPACE = -5

STENCIL = {"CIC_REF": 0, "CIC": 1, "TSC": 2}
WRAP_VEC_TRUE, WRAP_VEC_FALSE, WRAP_SCAN_TRUE, WRAP_SCAN_FALSE = 0, 1, 2, 3
FLAG_ZERO_DC = 1
FLAG_KMEAN = 2
STATS_LEN = 4
NUM_STAGES = 15
NOALLOC_SCRATCH_BYTES = 16 << 20

# (name, restype, argtypes) for every symbol declared in include/pk_api.h
_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_f = C.c_float
_d = C.c_double
_sz = C.c_size_t
_psz = C.POINTER(C.c_size_t)
XI_MAX_BINS = 1024
BISPEC_MAX_SHELLS = 1024
XI_EMPTY_NAN = 4
FIELD_DELTA = 0
FIELD_PSI_Z = 1
BATCH_GRAPH = 1
BATCH_CAPACITY_GRIDS = 2


class PkJob(C.Structure):
    """struct pk_job of pk_api.h (one catalogue of a batch)."""
    _fields_ = [
        ("x", _vp), ("sx", _i64), ("y", _vp), ("sy", _i64), ("z", _vp), ("sz", _i64),
        ("w", _vp), ("sw", _i64), ("np", _i64),
        ("xmin", _f), ("ymin", _f), ("zmin", _f), ("box_size", _d),
        ("stencil", _i), ("wrap_mode", _i), ("window_p", _i), ("flags", _i),
        ("thr", _vp), ("jstart", _vp), ("modes", _vp), ("ksum", _vp), ("nb", _i), ("shot_noise", _d),
        ("sums", _vp), ("counts", _vp), ("pk", _vp), ("nmodes", _vp), ("kmean", _vp), ("stats", _vp),
    ]


SIGNATURES = [
    ("pk_version", _i, []),
    ("pk_last_error", C.c_char_p, []),
    ("pk_plan_create", _i, [_i, C.POINTER(_vp)]),
    ("pk_plan_destroy", _i, [_vp]),
    ("pk_supported_n", _i, [_i]),
    ("pk_scatter_workspace_size", _i, [_i, _i64, _i, _vp, _psz]),
    ("pk_scatter", _i, [_vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _f, _f, _f, _d, _i, _i, _i, _vp,
                        _vp, _vp, _vp, _vp, _sz, _vp]),
    ("pk_r2c", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("pk_powspec_workspace_size", _i, [_vp, _i, _psz]),
    ("pk_powspec_mesh", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp,
                             _vp, _vp, _sz, _vp]),
    ("pk_powspec_mesh_noalloc", _i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _vp, _vp, _vp, _vp,
                                     _vp, _vp, _vp]),
    ("pk_estimate_workspace_size", _i, [_vp, _i64, _i, _vp, _i, _psz]),
    ("pk_estimate", _i, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _f, _f, _f, _d, _i, _i, _vp,
                         _vp, _vp, _vp, _vp, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    ("pk_estimate_backward_workspace_size", _i, [_vp, _i64, _i, _psz]),
    ("pk_estimate_backward", _i, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _f, _f, _f, _d, _i, _i, _vp,
                                  _vp, _vp, _i, _i, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    ("pk_batch_workspace_size", _i, [_vp, _i, _i64, _i, _vp, _psz]),
    ("pk_batch_create", _i, [_vp, _i, _i64, _i, _vp, _i, _vp, _sz, C.POINTER(_vp)]),
    ("pk_batch_set_jobs", _i, [_vp, _i, _vp, _vp]),
    ("pk_batch_run", _i, [_vp, _vp]),
    ("pk_batch_tile", _i, [_vp, _vp]),
    ("pk_batch_set_tuning", _i, [_vp, _i, _i, _i]),
    ("pk_batch_destroy", _i, [_vp]),
    ("pk_bispec_workspace_size", _i, [_vp, _i, _psz]),
    ("pk_bispec_mesh", _i, [_vp, _vp, _vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    ("pk_xi_workspace_size", _i, [_vp, _i, _psz]),
    ("pk_xi_mesh", _i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    ("pk_gaussian_workspace_size", _i, [_i, _psz]),
    ("pk_gaussian_field", _i, [_vp, _d, _vp, _vp, _i, _i, C.c_uint64, _vp, _vp, _sz, _vp]),
    ("pk_c2r", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("pk_gaussian_mesh", _i, [_vp, _d, _vp, _vp, _i, _i, C.c_uint64, _i, _vp, _vp, _vp, _sz, _vp]),
    ("pk_populate_workspace_size", _i, [_i, _psz]),
    ("pk_populate_count", _i, [_vp, _i, _d, _d, _d, C.c_uint64, _vp, _sz, _vp, _vp]),
    ("pk_populate_place", _i, [_vp, _vp, _d, _i, _d, _d, C.c_uint64, _vp, _sz, _vp, _i64, _vp]),
    ("pk_timing_enable", _i, [_i]),
    ("pk_timing_select", _i, [_i]),
    ("pk_timing_read", _i, [_vp, _vp, _i]),
    ("pk_stage_name", C.c_char_p, [_i]),
]

_lock = threading.Lock()
_lib = None


class PkError(RuntimeError):
    """A libpowspec_amd call returned an error status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def load(path: str | None = None):
    """Load the shared library (once).  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = path or _build.lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"HIP library {path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(no CPU fallback exists for the power-spectrum path)")
        lib = C.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(fn: str, code: int):
    if code == PK_OK:
        return
    msg = load().pk_last_error().decode(errors="replace")
    if code == PK_E_INVALID:
        raise ValueError(f"{fn}: {msg}")
    raise PkError(fn, code, msg)


def call(fn: str, *args):
    lib = load()
    check(fn, getattr(lib, fn)(*args))


def tile_arg(tile):
    """None -> NULL (library default); (tx, ty, tz) -> int[3]; 'direct' -> {0,0,0}."""
    if tile is None:
        return None, None
    if tile == "direct":
        tile = (0, 0, 0)
    arr = (C.c_int * 3)(*[int(v) for v in tile])
    return C.cast(arr, C.c_void_p), arr   # keep `arr` alive alongside the pointer
