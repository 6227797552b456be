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


def cache_key(tris, material, eps, *, layout_tag: str = SOURCE_MAJOR, **kwargs) -> str:
    """A 16-hex-digit key of everything the interaction matrix depends on:
    the vertex array itself (SHA-256 of its float64 bytes, so a reordered
    or reversed mesh is a different key), the material, eps (scalar or
    per-source), every keyword, the layout tag and this library's ABI
    version (studies/active_fractal/kernel.py:131-155 keys on the same
    things)."""
    tris = np.ascontiguousarray(np.asarray(tris, float))
    payload = {
        "geometry_sha256": hashlib.sha256(tris.tobytes()).hexdigest(),
        "n": int(tris.reshape(-1, 3, 3).shape[0]),
        "mu": float(material.mu), "lam": float(material.lam),
        "eps": _digest(np.asarray(eps, float)) if np.ndim(eps) else float(eps),
        "kwargs": {k: _digest(v) for k, v in sorted(kwargs.items())},
        "layout": layout_tag,
        "abi": _lib.ABI_VERSION,
    }
    blob = json.dumps(payload, sort_keys=True, default=str).encode()
    return hashlib.sha256(blob).hexdigest()[:16]
