"""The public matrix functions: the drop-in for src/mhs/matrices.py.

Same names, signatures (pinned by tests/gates/mhs/verify_import_hygiene.py:
159-180), output shapes, conventions, validation order, exception types and
byte ceiling as the reference; the per-source assembly loops below
src/mhs/kernels/assemble.py are replaced by the HIP kernels (mhs_amd.runtime).
"""
from __future__ import annotations

import threading

import numpy as np

from . import chunking, defaults, runtime
from .materials import Material
from .shared import CONTINUOUS_ORDER, SharedNodes, check_layout, shared_nodes

__all__ = ["disp_matrix", "stress_matrix", "total_stress_matrix", "eigenstress_matrix",
           "elastic_strain_matrix", "traction_matrix", "interaction_matrix",
           "collocation_points"]

COLLOCATION_SHRINK = 0.5


def _coerce(obs_pts, tris, material, eps):
    """matrices.py:75-128."""
    if not isinstance(material, Material):
        raise TypeError(
            f"material must be an mhs_amd.Material(mu, lam), got {type(material).__name__}. "
            f"Use Material.from_mu_nu(mu, nu) to convert someone else's parameters once.")
    obs = np.ascontiguousarray(np.asarray(obs_pts, dtype=np.float64))
    tri = np.ascontiguousarray(np.asarray(tris, dtype=np.float64))
    if obs.ndim != 2 or obs.shape[1] != 3:
        raise ValueError(f"obs_pts must have shape (n_obs, 3), got {obs.shape}")
    if tri.ndim != 3 or tri.shape[1:] != (3, 3):
        raise ValueError(f"tris must have shape (n_src, 3, 3) = (triangle, vertex, xyz), got {tri.shape}")
    if obs.shape[0] == 0 or tri.shape[0] == 0:
        raise ValueError(f"nothing to assemble: {obs.shape[0]} obs points, {tri.shape[0]} source triangles")
    if not np.all(np.isfinite(obs)) or not np.all(np.isfinite(tri)):
        raise ValueError("obs_pts and tris must be finite")
    if np.any(obs[:, 2] > 0.0):
        bad = int(np.argmax(obs[:, 2]))
        raise ValueError(f"the half space is z <= 0 with the free surface at z = 0, but "
                         f"obs_pts[{bad}] has z = {obs[bad, 2]:+.6g} > 0")
    if np.any(tri[:, :, 2] > 0.0):
        s, v = np.unravel_index(int(np.argmax(tri[:, :, 2])), tri.shape[:2])
        raise ValueError(f"the half space is z <= 0, but tris[{s}] vertex {v} has "
                         f"z = {tri[s, v, 2]:+.6g} > 0")
    eps_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (tri.shape[0],)))
    if not np.all(np.isfinite(eps_arr)) or np.any(eps_arr <= 0.0):
        raise ValueError("eps must be finite and strictly positive -- one value, or one per source "
                         "triangle. eps = 0 is the classical singular kernel, which is cutde's job.")
    return obs, tri, float(material.mu), float(material.lam), eps_arr


def _order(order) -> int:
    p = int(order)
    if p not in (0, 1, 2):
        raise ValueError(f"order must be 0, 1 or 2 (P0, P1, P2); got {order!r}")
    return p


def _coerce_normals(normals, n_obs):
    arr = np.atleast_2d(np.asarray(normals, float))
    if arr.shape == (1, 3) and n_obs != 1:
        arr = np.repeat(arr, n_obs, axis=0)
    if arr.shape != (n_obs, 3):
        raise ValueError(f"obs_normals has shape {np.shape(normals)}; expected ({n_obs}, 3) to "
                         f"match obs_pts, or (3,) for one shared normal")
    mag = np.linalg.norm(arr, axis=1)
    if not np.all(mag > 1e-13):
        raise ValueError(f"obs_normals[{int(np.argmin(mag))}] has length {float(mag.min()):.3e} "
                         f"and so names no plane")
    return arr / mag[:, None]


def _prepare_out(out, shape, n_obs, n_src, kind, parts):
    """matrices.py:169-189: validate out=, or budget-check the allocation
    (the host ceiling; a device-resident plan, :func:`plan_device`, is
    admitted against the device's memory by its caller instead)."""
    if out is None:
        if not getattr(_tls, "device_only", False):
            chunking.refuse_if_too_large(n_obs, n_src, kind, np.float64, parts)
        return None
    if not isinstance(out, np.ndarray):
        raise TypeError(f"out must be a numpy array, got {type(out).__name__}")
    if out.shape != shape:
        raise ValueError(f"out has shape {out.shape}, expected {shape}")
    if out.dtype not in (np.float64, np.float32):
        raise ValueError(f"out must be float64 or float32, got {out.dtype}")
    return out


def _columns(tri, p, layout, shared):
    """(ids, n_dof, shared) (matrices.py:192-220)."""
    check_layout(layout)
    if layout == "discontinuous":
        if shared is not None:
            raise ValueError("shared= is the continuous layout's table; pass layout='continuous' with it")
        return None, tri.shape[0] * (p + 1) * (p + 2) // 2, None
    if p != CONTINUOUS_ORDER:
        raise ValueError(f"layout='continuous' is quadratic only (order={CONTINUOUS_ORDER}), got order={p}")
    sn = shared_nodes(tri) if shared is None else shared
    if not isinstance(sn, SharedNodes):
        raise TypeError(f"shared must be a SharedNodes from shared_nodes, got {type(sn).__name__}")
    if sn.n_tri != tri.shape[0]:
        raise ValueError(f"shared describes {sn.n_tri} triangles, tris has {tri.shape[0]}")
    return sn.ids, sn.n, sn


_tls = threading.local()


def _finish(job, buf):
    """Run on the GPU and land the result in the caller's buffer or a new one.
    While a planning capture is open on this thread (:func:`plan`), record the
    validated job instead and return None."""
    cap = getattr(_tls, "capture", None)
    if cap is not None:
        cap.append(job)
        return None
    if buf is not None and not job.parts and 8 * buf.size > defaults.DEVICE_BLOCK_BYTES:
        # matrices.py:169-189: out= may be larger than memory (a memmap);
        # it is filled by receiver-row blocks of one device block each
        from . import parallel
        return parallel.stream_job(job, buf)
    out, out2 = runtime.run(job)
    res = runtime.to_host(out, buf)
    if job.parts:
        return res, runtime.to_host(out2)
    return res


def _job(kind, obs, tri, mu, lam, eps_arr, p, ids, sn, fold, **kw):
    return runtime.Job(kind=kind, order=p, obs=obs, tris=tri, eps=eps_arr, mu=mu, lam=lam,
                       fold=bool(fold), ids=ids, n_nodes=(sn.n if sn is not None else 0),
                       mixed=runtime.get_precision() == "mixed", **kw)


def disp_matrix(obs_pts, tris, material, eps, *, order=0, layout="discontinuous", shared=None,
                out=None, fold=True):
    """Displacement matrix (n_obs, 3, n_dof, 3) (matrices.py:231-248)."""
    obs, tri, mu, lam, eps_arr = _coerce(obs_pts, tris, material, eps)
    p = _order(order)
    ids, n_dof, sn = _columns(tri, p, layout, shared)
    buf = _prepare_out(out, (obs.shape[0], 3, n_dof, 3), obs.shape[0], n_dof, "disp", False)
    return _finish(_job("disp", obs, tri, mu, lam, eps_arr, p, ids, sn, fold), buf)


def stress_matrix(obs_pts, tris, material, eps, *, subtract_eigenstress=True, parts=False, order=0,
                  layout="discontinuous", shared=None, out=None, fold=True):
    """ELASTIC stress (n_obs, 3, 3, n_dof, 3); parts=True -> (stress, eigenstress)
    (matrices.py:251-295)."""
    obs, tri, mu, lam, eps_arr = _coerce(obs_pts, tris, material, eps)
    p = _order(order)
    ids, n_dof, sn = _columns(tri, p, layout, shared)
    shape = (obs.shape[0], 3, 3, n_dof, 3)
    buf = _prepare_out(out, shape, obs.shape[0], n_dof, "stress", parts)
    if parts and out is not None:
        raise ValueError("parts=True returns two arrays, so out= cannot name the destination of "
                         "both; call twice with out=, or drop out=.")
    return _finish(_job("stress", obs, tri, mu, lam, eps_arr, p, ids, sn, fold,
                        subtract=bool(subtract_eigenstress), parts=bool(parts)), buf)


def traction_matrix(obs_pts, obs_normals, tris, material, eps, *, subtract_eigenstress=True, order=0,
                    layout="discontinuous", shared=None, out=None, fold=True):
    """ELASTIC traction (n_obs, 3, n_dof, 3) (matrices.py:298-329)."""
    obs, tri, mu, lam, eps_arr = _coerce(obs_pts, tris, material, eps)
    p = _order(order)
    ids, n_dof, sn = _columns(tri, p, layout, shared)
    nrm = _coerce_normals(obs_normals, obs.shape[0])
    buf = _prepare_out(out, (obs.shape[0], 3, n_dof, 3), obs.shape[0], n_dof, "traction", False)
    return _finish(_job("traction", obs, tri, mu, lam, eps_arr, p, ids, sn, fold,
                        subtract=bool(subtract_eigenstress), nrm=nrm), buf)


def collocation_points(tris, order: int = 0, shrink: float = COLLOCATION_SHRINK,
                       layout: str = "discontinuous", shared=None) -> np.ndarray:
    """(n_dof, 3) collocation points (matrices.py:348-375)."""
    tri = np.asarray(tris, float).reshape(-1, 3, 3)
    p = _order(order)
    _, _, sn = _columns(tri, p, layout, shared)
    if sn is not None:
        if shrink != COLLOCATION_SHRINK:
            raise ValueError("shrink does not apply to the continuous layout: its collocation "
                             "points are the shared nodes")
        return sn.coords.copy()
    if p == 0:
        bary = np.full((1, 3), 1.0 / 3.0)
    elif p == 1:
        bary = np.eye(3)
    else:
        bary = np.array([[2, 0, 0], [0, 2, 0], [0, 0, 2], [1, 1, 0], [0, 1, 1], [1, 0, 1]]) / 2.0
    bary = bary + float(shrink) * (1.0 / 3.0 - bary)
    return (bary[None, :, :] @ tri).reshape(-1, 3)


def _rows(rows, n_rec):
    if rows is None:
        return 0, n_rec
    lo, hi = int(rows[0]), int(rows[1])
    if not (0 <= lo < hi <= n_rec):
        raise ValueError(f"rows={rows!r} is not a block of the {n_rec} receiver rows")
    return lo, hi


def interaction_matrix(tris, material, eps, *, receiver=("strike", "dip", "normal"),
                       source=("strike", "dip", "tensile"), obs_tris=None, obs_pts=None, order=0,
                       shrink=COLLOCATION_SHRINK, subtract_eigenstress=True, layout="discontinuous",
                       shared=None, rows=None, out=None, fold=True):
    """(n_rec_dof, n_src_dof, n_a, n_b) on-fault interaction (matrices.py:378-491)."""
    from . import tdcs
    tri = np.asarray(tris, float).reshape(-1, 3, 3)
    p = _order(order)
    K = (p + 1) * (p + 2) // 2
    ids, n_src_dof, sn = _columns(tri, p, layout, shared)
    if sn is not None:
        if obs_tris is not None or obs_pts is not None:
            raise ValueError("layout='continuous' collocates at the shared nodes of tris: obs_tris "
                             "and obs_pts cannot be given.")
        if shrink != COLLOCATION_SHRINK:
            raise ValueError("shrink does not apply to the continuous layout: its collocation "
                             "points are the shared nodes")
        pts = sn.coords
        rec_frame = sn.frames
        src_frame = sn.frames
        n_rec_dof = sn.n
    else:
        rec_tri = tri if obs_tris is None else np.asarray(obs_tris, float).reshape(-1, 3, 3)
        pts = collocation_points(rec_tri, p, shrink) if obs_pts is None else obs_pts
        n_rec_dof = rec_tri.shape[0] * K
        if np.asarray(pts).shape[0] != n_rec_dof:
            raise ValueError(
                f"interaction_matrix pairs observers with RECEIVER slip DOFs row by row, so at order "
                f"{p} it needs {K} per receiver element: got {np.asarray(pts).shape[0]} observers for "
                f"{rec_tri.shape[0]} receiver triangles ({n_rec_dof} DOFs).")
        rec_frame = np.repeat(tdcs.slip_frame(rec_tri), K, axis=0)
        src_frame = tdcs.slip_frame(tri)
    lo, hi = _rows(rows, n_rec_dof)
    obs, tri, mu, lam, eps_arr = _coerce(np.asarray(pts, float)[lo:hi], tri, material, eps)
    rec = tdcs.component_index(receiver)
    src = tdcs.component_index(source)
    shape = (hi - lo, n_src_dof, len(rec), len(src))
    buf = _prepare_out(out, shape, hi - lo, n_src_dof, (len(rec), len(src)), False)
    job = _job("interaction", obs, tri, mu, lam, eps_arr, p, ids, sn, fold,
               subtract=bool(subtract_eigenstress),
               self_rows=sn is not None or (obs_tris is None and obs_pts is None),
               nrm=np.ascontiguousarray(rec_frame[lo:hi, 2, :]),
               rb=np.ascontiguousarray(rec_frame[lo:hi][:, rec, :]),
               sb=np.ascontiguousarray(src_frame[:, src, :]))
    return _finish(job, buf)


def total_stress_matrix(obs_pts, tris, material, eps, *, order=0, layout="discontinuous", shared=None,
                        out=None, fold=True):
    """RAW mollified stress: elastic + eigenstress (matrices.py:494-515)."""
    obs, tri, mu, lam, eps_arr = _coerce(obs_pts, tris, material, eps)
    p = _order(order)
    ids, n_dof, sn = _columns(tri, p, layout, shared)
    buf = _prepare_out(out, (obs.shape[0], 3, 3, n_dof, 3), obs.shape[0], n_dof, "stress", False)
    return _finish(_job("stress", obs, tri, mu, lam, eps_arr, p, ids, sn, fold, subtract=False), buf)


def eigenstress_matrix(obs_pts, tris, material, eps, *, order=0, layout="discontinuous", shared=None,
                       out=None, fold=True):
    """The eigenstress C:eps* alone (matrices.py:518-543)."""
    obs, tri, mu, lam, eps_arr = _coerce(obs_pts, tris, material, eps)
    p = _order(order)
    ids, n_dof, sn = _columns(tri, p, layout, shared)
    buf = _prepare_out(out, (obs.shape[0], 3, 3, n_dof, 3), obs.shape[0], n_dof, "eigenstress", False)
    return _finish(_job("eigen", obs, tri, mu, lam, eps_arr, p, ids, sn, fold), buf)


def elastic_strain_matrix(obs_pts, tris, material, eps, *, subtract_eigenstress=True, order=0,
                          layout="discontinuous", shared=None, out=None, fold=True):
    """ELASTIC strain as a full (3, 3) tensor (matrices.py:546-580)."""
    obs, tri, mu, lam, eps_arr = _coerce(obs_pts, tris, material, eps)
    p = _order(order)
    ids, n_dof, sn = _columns(tri, p, layout, shared)
    buf = _prepare_out(out, (obs.shape[0], 3, 3, n_dof, 3), obs.shape[0], n_dof, "strain", False)
    return _finish(_job("stress", obs, tri, mu, lam, eps_arr, p, ids, sn, fold,
                        subtract=bool(subtract_eigenstress), strain=True), buf)


def plan(fn, *args, **kwargs) -> runtime.Job:
    """The validated :class:`runtime.Job` that ``fn(*args, **kwargs)`` (one of
    the public matrix functions) would run, without running it: the same
    coercion, checks and byte ceiling, for callers that execute jobs
    themselves (the serving loop)."""
    _tls.capture = []
    try:
        fn(*args, **kwargs)
        jobs = _tls.capture
    finally:
        _tls.capture = None
    if len(jobs) != 1:
        raise ValueError(f"{getattr(fn, '__name__', fn)!r} is not a single-job matrix function")
    return jobs[0]


def plan_device(fn, *args, **kwargs) -> runtime.Job:
    """:func:`plan` for a result that stays in device memory (the serving
    loop): the same coercion and checks, but not the host's byte ceiling
    (chunking.MATRIX_MAX_BYTES bounds host RAM; the caller admits the job
    against the device's memory, runtime.job_bytes)."""
    _tls.device_only = True
    try:
        return plan(fn, *args, **kwargs)
    finally:
        _tls.device_only = False

