"""ctypes binding of the C-ABI in include/mhs_amd.h (libmhs_amd.so).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m mhs_amd.build``).  Loading fails loudly when it is missing: there
is no other implementation to fall back to.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libmhs_amd.so"
ABI_VERSION = 6

KIND = {"disp": 0, "stress": 1, "eigen": 2, "traction": 3, "interaction": 4}
F_FOLD, F_SUBTRACT, F_PARTS, F_MIXED = 1, 2, 4, 8
REFUSE_DEGENERATE, REFUSE_NON_MANIFOLD, REFUSE_SHARED, REFUSE_WINDING = 1, 2, 4, 8
STAGE_PAIR, STAGE_Q, STAGE_FOLD = 1, 2, 4


RULE_BANDS = 8


class Rule(C.Structure):
    """mirror of ``mhs_rule`` (the mixed mode's far-field rule; near_ratio 0:
    the library's default)"""
    _fields_ = [("near_ratio", C.c_double), ("band_hi", C.c_double * RULE_BANDS),
                ("order", C.c_int32 * RULE_BANDS), ("fold_near_ratio", C.c_double)]


class Launch(C.Structure):
    """mirror of ``mhs_launch``"""
    _fields_ = [
        ("abi_version", C.c_int32), ("kind", C.c_int32), ("order", C.c_int32),
        ("flags", C.c_int32), ("n_a", C.c_int32), ("n_b", C.c_int32),
        ("n_obs", C.c_int64), ("n_src", C.c_int64), ("n_cols", C.c_int64),
        ("obs_lo", C.c_int64), ("obs_hi", C.c_int64),
        ("mu", C.c_double), ("lam", C.c_double),
        ("obs", C.c_void_p), ("nrm", C.c_void_p), ("rb", C.c_void_p), ("sb", C.c_void_p),
        ("src", C.c_void_p), ("coef", C.c_void_p), ("gl_x", C.c_void_p), ("gl_w", C.c_void_p),
        ("gq", C.c_void_p),
        ("out", C.c_void_p), ("out2", C.c_void_p),
        ("strad", C.c_void_p), ("n_strad", C.c_int64),
        ("fold_list", C.c_void_p), ("fold_count", C.c_void_p), ("fold_err", C.c_void_p),
        ("fold_cap", C.c_int64), ("stages", C.c_int32), ("far_tiles", C.c_int32),
        ("fold_work", C.c_void_p), ("fold_work_items", C.c_int64),
        ("rule", Rule), ("near_tiles", C.c_void_p),
    ]


_LIB = None

EXPORTS = ("mhs_assemble", "mhs_histogram", "mhs_gather", "mhs_strain", "mhs_fill_material", "mhs_ncoef",
           "mhs_src_stride", "mhs_abi_version", "mhs_last_error", "mhs_fold_item_bytes", "mhs_source_major",
           "mhs_subtract", "mhs_assemble_static", "mhs_batch_workspace_bytes", "mhs_assemble_batch",
           "mhs_fold_histogram",
           "mhs_weld_workspace_bytes", "mhs_shared_nodes", "mhs_tri_geometry")


def load(path=None):
    """Load and type the library; raise if absent or of another ABI."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = pathlib.Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first "
            f"(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    lib.mhs_assemble.argtypes = [C.POINTER(Launch), C.c_void_p]
    lib.mhs_assemble.restype = C.c_int
    lib.mhs_assemble_static.argtypes = [C.POINTER(Launch), C.c_void_p]
    lib.mhs_assemble_static.restype = C.c_int
    lib.mhs_batch_workspace_bytes.argtypes = [C.POINTER(Launch), C.c_int]
    lib.mhs_batch_workspace_bytes.restype = C.c_int64
    lib.mhs_assemble_batch.argtypes = [C.POINTER(Launch), C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    lib.mhs_assemble_batch.restype = C.c_int
    lib.mhs_histogram.argtypes = [C.POINTER(Launch), C.c_void_p, C.c_void_p]
    lib.mhs_fold_histogram.argtypes = [C.POINTER(Launch), C.c_void_p, C.c_void_p]
    lib.mhs_fold_histogram.restype = C.c_int
    lib.mhs_histogram.restype = C.c_int
    lib.mhs_gather.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.mhs_gather.restype = C.c_int
    lib.mhs_strain.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_void_p]
    lib.mhs_strain.restype = C.c_int
    lib.mhs_subtract.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    lib.mhs_subtract.restype = C.c_int
    lib.mhs_fill_material.argtypes = [C.c_double, C.c_double, C.c_void_p]
    lib.mhs_fill_material.restype = C.c_int
    lib.mhs_source_major.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_void_p]
    lib.mhs_source_major.restype = C.c_int
    lib.mhs_weld_workspace_bytes.argtypes = [C.c_int64]
    lib.mhs_weld_workspace_bytes.restype = C.c_int64
    lib.mhs_shared_nodes.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int64] + [C.c_void_p] * 9 + [C.POINTER(C.c_int64),
                                                                                  C.POINTER(C.c_int),
                                                                                  C.c_void_p]
    lib.mhs_tri_geometry.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double), C.c_void_p]
    lib.mhs_tri_geometry.restype = C.c_int
    lib.mhs_shared_nodes.restype = C.c_int
    lib.mhs_fold_item_bytes.argtypes = [C.c_int]
    lib.mhs_fold_item_bytes.restype = C.c_int64
    for nm in ("mhs_ncoef", "mhs_src_stride", "mhs_abi_version"):
        getattr(lib, nm).restype = C.c_int
    lib.mhs_last_error.restype = C.c_char_p
    if lib.mhs_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{p} has ABI {lib.mhs_abi_version()}, expected {ABI_VERSION}")
    if path is None:
        _LIB = lib
    return lib


def check(rc: int, lib=None) -> None:
    if rc != 0:
        lib = lib or load()
        raise RuntimeError(f"mhs_amd: {lib.mhs_last_error().decode(errors='replace')}")
