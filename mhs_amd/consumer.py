"""Consumer-side forms of the interaction matrix.

Quasi-static cycle and cascade codes read the interaction matrix one SOURCE
at a time ("element j slipped: update the traction on every receiver"), so
they store it source-major, ``(n_src, n_b, n_a, n_rec)``, and key an on-disk
cache by a hash of the vertices themselves
(studies/active_fractal/kernel.py:14-23, 131-155, 259-261).  Here the
transpose runs on the device right after the assembly (``mhs_source_major``,
an HBM-bound tiled copy) instead of as a strided numpy transpose of the whole
matrix on the host, and :func:`cache_key` gives the content key.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json

import numpy as np

from . import _lib, api, runtime

#: The layout tag folded into :func:`cache_key`.
SOURCE_MAJOR = "source_major_(n_src,n_b,n_a,n_rec)"


def interaction_source_major(tris, material, eps, *, out=None, device: bool = False, **kwargs):
    """``interaction_matrix(tris, material, eps, **kwargs)`` in the
    source-major layout: a contiguous ``(n_src_dof, n_b, n_a, n_rec_dof)``
    array with ``result[s, b, a, r] == interaction_matrix(...)[r, s, a, b]``
    (bitwise: the same device assembly, then a copy).

    Every keyword of ``interaction_matrix`` is accepted except its ``out=``;
    ``out`` here is a float64 or float32 ndarray of the source-major shape.
    ``device=True`` returns the device tensor instead of a host array.
    Validation and errors are ``interaction_matrix``'s."""
    job = api.plan(api.interaction_matrix, tris, material, eps, **kwargs)
    n_rec, n_src, n_a, n_b = job.out_shape()
    shape = (n_src, n_b, n_a, n_rec)
    if out is not None:                  # refused before any device work
        if not isinstance(out, np.ndarray):
            raise TypeError(f"out must be a numpy.ndarray, got {type(out).__name__}")
        if out.shape != shape:
            raise ValueError(f"out has shape {out.shape}, expected {shape}")
        if out.dtype not in (np.float64, np.float32):
            raise ValueError(f"out must be float64 or float32, got {out.dtype}")
    prep = runtime.Prepared(job)
    torch = prep.dev.torch
    with torch.cuda.stream(prep.stream):
        prep.launch()
    prep.check_fold()
    K = prep.final
    dst = torch.empty(shape, dtype=torch.float64, device=K.device)
    lib = prep.dev.lib
    with torch.cuda.stream(prep.stream):
        _lib.check(lib.mhs_source_major(C.c_void_p(K.data_ptr()), C.c_void_p(dst.data_ptr()), n_rec, n_src,
                                        n_a, n_b, prep.handle), lib)
    prep.stream.synchronize()
    del K
    if device:
        return dst
    return runtime.to_host(dst, out)


def _digest(v) -> object:
    """JSON-able stand-in of a keyword value: arrays by their SHA-256."""
    if isinstance(v, np.ndarray):
        a = np.ascontiguousarray(v)
        return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "shape": list(a.shape), "dtype": str(a.dtype)}
    if isinstance(v, (tuple, list)):
        return [_digest(x) for x in v]
    if hasattr(v, "ids") and hasattr(v, "coords"):        # SharedNodes
        return {"shared_ids": _digest(np.asarray(v.ids)), "shared_coords": _digest(np.asarray(v.coords))}
    return v


#: interaction_matrix's keyword defaults (matrices.py:378-382): a keyword
#: left out and the same keyword given at its default are one matrix.
_DEFAULTS = {"receiver": ("strike", "dip", "normal"), "source": ("strike", "dip", "tensile"),
             "order": 0, "shrink": api.COLLOCATION_SHRINK, "subtract_eigenstress": True,
             "layout": "discontinuous", "fold": True}


def _normalised(kwargs: dict) -> dict:
    """The keywords with defaults filled in and equivalent spellings made
    one: components as frame-row indices ('normal' is 'tensile', a string
    is a one-element tuple), order as an int, booleans as bools."""
    from . import tdcs
    kw = dict(_DEFAULTS)
    kw.update(kwargs)
    for k in ("receiver", "source"):
        kw[k] = [int(i) for i in tdcs.component_index(kw[k])]
    kw["order"] = api._order(kw["order"])
    kw["shrink"] = float(kw["shrink"])
    for k in ("subtract_eigenstress", "fold"):
        kw[k] = bool(kw[k])
    return kw


def kernel_tag() -> dict:
    """What the device kernels do beyond the mesh and keywords, for the
    cache key: the precision mode in force (a matrix built with FP32
    far-field points is a different matrix from the FP64 one), the mixed
    far-field rule fixed by the caller (None: tuned or the library default,
    both bounded by the mixed mode's error bound), the near/far switch and
    far-field Gauss orders, the image Q-family law and the fold's constants
    (as studies/active_fractal/kernel.py:109-119 tags its fold)."""
    from . import defaults as d
    prec = runtime.get_precision()
    return {
        "precision": prec,
        "mixed_rule": (repr(runtime.mixed_rule) if prec == "mixed" else None),
        "near_far": [d.D_STAR, d.D_STAR_DISTANT, d.FAR_GAUSS_N, d.FAR_GAUSS_N_DISTANT],
        "image_q": [d.IMAGE_Q_BUDGET_C, d.IMAGE_Q_GAUSS_MIN, d.IMAGE_Q_GAUSS_MAX],
        "fold": [d.FOLD_SOURCE_DEPTH_EPS, d.FOLD_BAND_DEPTH_EPS, d.FOLD_REACH_EPS,
                 d.FOLD_FINE_SPACING_EPS, d.FOLD_N_RADIAL, d.FOLD_N_ANGULAR],
    }


def cache_key(tris, material, eps, *, layout_tag: str = SOURCE_MAJOR, **kwargs) -> str:
    """A 16-hex-digit key of everything the interaction matrix depends on:
    the vertex array itself (SHA-256 of its float64 bytes, so a reordered
    or reversed mesh is a different key), the material, eps (scalar or
    per-source), every keyword (defaults filled in, component names
    normalised), the layout tag, the kernel tag (:func:`kernel_tag`: the
    precision mode and rule in force, the thresholds) and this library's
    ABI version (studies/active_fractal/kernel.py:131-155 keys on the same
    things).  Call it under the precision the matrix is assembled in."""
    tris = np.ascontiguousarray(np.asarray(tris, float))
    payload = {
        "geometry_sha256": hashlib.sha256(tris.tobytes()).hexdigest(),
        "n": int(tris.reshape(-1, 3, 3).shape[0]),
        "mu": float(material.mu), "lam": float(material.lam),
        "eps": _digest(np.asarray(eps, float)) if np.ndim(eps) else float(eps),
        "kwargs": {k: _digest(v) for k, v in sorted(_normalised(kwargs).items())},
        "layout": layout_tag,
        "kernel": kernel_tag(),
        "abi": _lib.ABI_VERSION,
    }
    blob = json.dumps(payload, sort_keys=True, default=str).encode()
    return hashlib.sha256(blob).hexdigest()[:16]
