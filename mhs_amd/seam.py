"""The reference's assembly seam, served by the device kernels.

The reference's public matrix functions validate their inputs and then call
one of five ``assemble_*`` functions (src/mhs/kernels/assemble.py:152-366)
that fill a caller-allocated ``out`` from contiguous float64 arrays and plain
floats, one source triangle at a time.  Those five are the designed seam for
a compiled replacement (assemble.py:18-21, 96-105).  This module keeps their
names, signatures, argument meaning and ``out`` semantics, and runs each as
one device assembly through the C-ABI (``runtime.Prepared`` ->
``mhs_assemble``):

* ``ids=None``: ``out`` is overwritten with the matrix (the reference writes
  every source's disjoint column slice);
* ``ids`` given (``ids[s, k]`` = the column element ``s``'s node ``k`` writes
  into): the matrix is ADDED into ``out`` (the reference's scatter-add; a
  buffer written this way starts at zero).  Shared columns are summed on the
  device in ascending element order, the order the reference's source loop
  visits them.

``out`` is a float64 or float32 ndarray (a memmap too) of the form's shape;
an ``out`` larger than one device block (``defaults.DEVICE_BLOCK_BYTES``)
without ``ids`` is filled by receiver-row blocks.  The arithmetic is the
precision mode in force (``mhs_amd.precision``).  There is no CPU path: a
missing GPU or library raises.
"""
from __future__ import annotations

import numpy as np

from . import defaults, runtime

__all__ = ["assemble_halfspace_disp", "assemble_halfspace_total_stress", "assemble_halfspace_traction",
           "assemble_interaction", "assemble_eigenstress"]


def _arr(name, a, shape_tail, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if a.ndim != 1 + len(shape_tail) or tuple(a.shape[1:]) != tuple(shape_tail) or (n is not None and a.shape[0] != n):
        want = ("n" if n is None else str(n),) + tuple(str(t) for t in shape_tail)
        raise ValueError(f"{name} must have shape ({', '.join(want)}), got {a.shape}")
    if not np.all(np.isfinite(a)):
        raise ValueError(f"{name} must be finite")
    return a


def _common(obs, tris, mu, lam, eps_arr, order):
    obs = _arr("obs", obs, (3,))
    tris = _arr("tris", tris, (3, 3))
    eps = np.ascontiguousarray(np.asarray(eps_arr, dtype=np.float64).reshape(-1))
    if eps.shape != (tris.shape[0],):
        raise ValueError(f"eps_arr must have one value per source triangle ({tris.shape[0]}), got {eps.shape}")
    if not np.all(np.isfinite(eps)) or np.any(eps <= 0.0):
        raise ValueError("eps_arr must be finite and strictly positive")
    p = int(order)
    if p not in (0, 1, 2):
        raise ValueError(f"order must be 0, 1 or 2, got {order!r}")
    if obs.shape[0] == 0 or tris.shape[0] == 0:
        raise ValueError(f"nothing to assemble: {obs.shape[0]} observers, {tris.shape[0]} sources")
    return obs, tris, float(mu), float(lam), eps, p


def _ids(ids, n_src, K, n_cols):
    if ids is None:
        return None
    a = np.ascontiguousarray(np.asarray(ids), dtype=np.int64)
    if a.shape != (n_src, K):
        raise ValueError(f"ids must have shape ({n_src}, {K}), got {a.shape}")
    if a.size and (a.min() < 0 or a.max() >= n_cols):
        raise ValueError(f"ids must name columns in [0, {n_cols}) of out")
    return a


def _check_out(out, shape):
    if not isinstance(out, np.ndarray):
        raise TypeError(f"out must be a numpy.ndarray, got {type(out).__name__}")
    if out.shape != tuple(shape):
        raise ValueError(f"out has shape {out.shape}, expected {tuple(shape)}")
    if out.dtype not in (np.float64, np.float32):
        raise ValueError(f"out must be float64 or float32, got {out.dtype}")


def _run(job, out):
    """Assemble ``job`` and write (no ids) or add (ids) it into ``out``."""
    if job.ids is None:
        if 8 * out.size > defaults.DEVICE_BLOCK_BYTES:
            from . import parallel
            return parallel.stream_job(job, out)
        dev_out, _ = runtime.run(job)
        runtime.to_host(dev_out, out)
        return out
    dev_out, _ = runtime.run(job)
    out += runtime.to_host(dev_out)
    return out


def _job(kind, obs, tris, mu, lam, eps, p, ids, n_cols, fold, **kw):
    return runtime.Job(kind=kind, order=p, obs=obs, tris=tris, eps=eps, mu=mu, lam=lam, fold=bool(fold),
                       ids=ids, n_nodes=(n_cols if ids is not None else 0),
                       mixed=runtime.get_precision() == "mixed", **kw)


def assemble_halfspace_disp(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True):
    """Half-space displacement into ``out`` (n_obs, 3, n_cols, 3)
    (assemble.py:186-206)."""
    obs, tris, mu, lam, eps, p = _common(obs, tris, mu, lam, eps_arr, order)
    K = (p + 1) * (p + 2) // 2
    n_cols = out.shape[2] if ids is not None and isinstance(out, np.ndarray) and out.ndim == 4 else tris.shape[0] * K
    _check_out(out, (obs.shape[0], 3, n_cols, 3))
    job = _job("disp", obs, tris, mu, lam, eps, p, _ids(ids, tris.shape[0], K, n_cols), n_cols, fold)
    return _run(job, out)


def assemble_halfspace_total_stress(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True):
    """Half-space TOTAL stress (eigenstress included) into ``out``
    (n_obs, 3, 3, n_cols, 3) (assemble.py:209-230)."""
    obs, tris, mu, lam, eps, p = _common(obs, tris, mu, lam, eps_arr, order)
    K = (p + 1) * (p + 2) // 2
    n_cols = out.shape[3] if ids is not None and isinstance(out, np.ndarray) and out.ndim == 5 else tris.shape[0] * K
    _check_out(out, (obs.shape[0], 3, 3, n_cols, 3))
    job = _job("stress", obs, tris, mu, lam, eps, p, _ids(ids, tris.shape[0], K, n_cols), n_cols, fold,
               subtract=False)
    return _run(job, out)


def assemble_halfspace_traction(obs, nrm, tris, mu, lam, eps_arr, out, subtract_eigenstress=True, order=0,
                                ids=None, fold=True):
    """Cartesian traction on the planes ``nrm`` into ``out``
    (n_obs, 3, n_cols, 3) (assemble.py:270-300).  ``nrm`` rows are used as
    given (the public calls normalise them)."""
    obs, tris, mu, lam, eps, p = _common(obs, tris, mu, lam, eps_arr, order)
    nrm = _arr("nrm", nrm, (3,), obs.shape[0])
    K = (p + 1) * (p + 2) // 2
    n_cols = out.shape[2] if ids is not None and isinstance(out, np.ndarray) and out.ndim == 4 else tris.shape[0] * K
    _check_out(out, (obs.shape[0], 3, n_cols, 3))
    job = _job("traction", obs, tris, mu, lam, eps, p, _ids(ids, tris.shape[0], K, n_cols), n_cols, fold,
               subtract=bool(subtract_eigenstress), nrm=nrm)
    return _run(job, out)


def assemble_interaction(obs, nrm, rec_basis, tris, src_basis, mu, lam, eps_arr, out, subtract_eigenstress=True,
                         order=0, ids=None, fold=True):
    """Frame-resolved interaction into ``out`` (n_obs, n_cols, n_a, n_b)
    (assemble.py:303-331): traction on the plane ``nrm[o]`` along
    ``rec_basis[o, a]`` from unit slip along ``src_basis[s, b]`` (per
    source, or with ``ids`` per column: (n_cols, n_b, 3))."""
    obs, tris, mu, lam, eps, p = _common(obs, tris, mu, lam, eps_arr, order)
    n_obs = obs.shape[0]
    nrm = _arr("nrm", nrm, (3,), n_obs)
    rb = np.ascontiguousarray(np.asarray(rec_basis, dtype=np.float64))
    if rb.ndim != 3 or rb.shape[0] != n_obs or rb.shape[2] != 3:
        raise ValueError(f"rec_basis must have shape ({n_obs}, n_a, 3), got {rb.shape}")
    sb = np.ascontiguousarray(np.asarray(src_basis, dtype=np.float64))
    K = (p + 1) * (p + 2) // 2
    n_cols = out.shape[1] if ids is not None and isinstance(out, np.ndarray) and out.ndim == 4 else tris.shape[0] * K
    n_rows_sb = n_cols if ids is not None else tris.shape[0]
    if sb.ndim != 3 or sb.shape[0] != n_rows_sb or sb.shape[2] != 3:
        raise ValueError(f"src_basis must have shape ({n_rows_sb}, n_b, 3), got {sb.shape}")
    _check_out(out, (n_obs, n_cols, rb.shape[1], sb.shape[1]))
    job = _job("interaction", obs, tris, mu, lam, eps, p, _ids(ids, tris.shape[0], K, n_cols), n_cols, fold,
               subtract=bool(subtract_eigenstress), nrm=nrm, rb=rb, sb=sb)
    return _run(job, out)


def assemble_eigenstress(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True):
    """The eigenstress C:eps* into ``out`` (n_obs, 3, 3, n_cols, 3)
    (assemble.py:334-366)."""
    obs, tris, mu, lam, eps, p = _common(obs, tris, mu, lam, eps_arr, order)
    K = (p + 1) * (p + 2) // 2
    n_cols = out.shape[3] if ids is not None and isinstance(out, np.ndarray) and out.ndim == 5 else tris.shape[0] * K
    _check_out(out, (obs.shape[0], 3, 3, n_cols, 3))
    job = _job("eigen", obs, tris, mu, lam, eps, p, _ids(ids, tris.shape[0], K, n_cols), n_cols, fold)
    return _run(job, out)
