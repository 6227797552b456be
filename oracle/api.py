"""The oracle's matrix forms, with the reference's signatures (matrices.py:231-580)
but (mu, lam) instead of a Material and no validation: callers pass what the
product validated.  Frames, collocation points and the shared-node table are
host logic pinned separately against the reference's outputs, and are taken
from mhs_amd's host layer.
"""
from __future__ import annotations

import numpy as np

from . import assemble as A
from .geometry import n_nodes


def _prep(obs, tris, eps):
    obs = np.ascontiguousarray(np.asarray(obs, float).reshape(-1, 3))
    tri = np.ascontiguousarray(np.asarray(tris, float).reshape(-1, 3, 3))
    eps_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, float), (tri.shape[0],)))
    return obs, tri, eps_arr


def _ids(tri, order, layout, shared):
    if layout == "discontinuous":
        return None, tri.shape[0] * n_nodes(order), None
    from mhs_amd.shared import shared_nodes
    sn = shared_nodes(tri) if shared is None else shared
    return sn.ids, sn.n, sn


def disp_matrix(obs, tris, mu, lam, eps, order=0, layout="discontinuous", shared=None, fold=True):
    obs, tri, e = _prep(obs, tris, eps)
    ids, n, _ = _ids(tri, order, layout, shared)
    out = np.zeros((obs.shape[0], 3, n, 3))
    return A.disp(obs, tri, mu, lam, e, out, order, ids, fold)


def total_stress_matrix(obs, tris, mu, lam, eps, order=0, layout="discontinuous", shared=None, fold=True):
    obs, tri, e = _prep(obs, tris, eps)
    ids, n, _ = _ids(tri, order, layout, shared)
    out = np.zeros((obs.shape[0], 3, 3, n, 3))
    return A.total_stress(obs, tri, mu, lam, e, out, order, ids, fold)


def eigenstress_matrix(obs, tris, mu, lam, eps, order=0, layout="discontinuous", shared=None, fold=True):
    obs, tri, e = _prep(obs, tris, eps)
    ids, n, _ = _ids(tri, order, layout, shared)
    out = np.zeros((obs.shape[0], 3, 3, n, 3))
    return A.eigenstress(obs, tri, mu, lam, e, out, order, ids, fold)


def stress_matrix(obs, tris, mu, lam, eps, subtract_eigenstress=True, parts=False, order=0,
                  layout="discontinuous", shared=None, fold=True):
    """matrices.py:251-295: total - eigenstress."""
    sig = total_stress_matrix(obs, tris, mu, lam, eps, order, layout, shared, fold)
    if not (subtract_eigenstress or parts):
        return sig
    eig = eigenstress_matrix(obs, tris, mu, lam, eps, order, layout, shared, fold)
    if subtract_eigenstress:
        sig = sig - eig
    return (sig, eig) if parts else sig


def elastic_strain_matrix(obs, tris, mu, lam, eps, subtract_eigenstress=True, order=0,
                          layout="discontinuous", shared=None, fold=True):
    """matrices.py:546-580."""
    buf = stress_matrix(obs, tris, mu, lam, eps, subtract_eigenstress, False, order, layout, shared, fold)
    tr = buf[:, 0, 0] + buf[:, 1, 1] + buf[:, 2, 2]
    tre = tr / (3.0 * lam + 2.0 * mu)
    for a in range(3):
        buf[:, a, a] -= lam * tre
    return buf / (2.0 * mu)


def traction_matrix(obs, normals, tris, mu, lam, eps, subtract_eigenstress=True, order=0,
                    layout="discontinuous", shared=None, fold=True):
    obs, tri, e = _prep(obs, tris, eps)
    nrm = np.asarray(normals, float).reshape(-1, 3)
    if nrm.shape[0] == 1 and obs.shape[0] != 1:
        nrm = np.repeat(nrm, obs.shape[0], axis=0)
    nrm = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    ids, n, _ = _ids(tri, order, layout, shared)
    out = np.zeros((obs.shape[0], 3, n, 3))
    return A.traction(obs, nrm, tri, mu, lam, e, out, subtract_eigenstress, order, ids, fold)


def interaction_matrix(tris, mu, lam, eps, receiver=("strike", "dip", "normal"),
                       source=("strike", "dip", "tensile"), obs_tris=None, obs_pts=None, order=0,
                       shrink=0.5, subtract_eigenstress=True, layout="discontinuous", shared=None,
                       rows=None, fold=True):
    """matrices.py:378-491."""
    from mhs_amd import tdcs
    from mhs_amd.api import collocation_points
    tri = np.asarray(tris, float).reshape(-1, 3, 3)
    K = n_nodes(order)
    ids, n_src_dof, sn = _ids(tri, order, layout, shared)
    if sn is not None:
        pts, rec_frame, src_frame, n_rec = sn.coords, sn.frames, sn.frames, sn.n
    else:
        rec_tri = tri if obs_tris is None else np.asarray(obs_tris, float).reshape(-1, 3, 3)
        pts = collocation_points(rec_tri, order, shrink) if obs_pts is None else obs_pts
        rec_frame = np.repeat(tdcs.slip_frame(rec_tri), K, axis=0)
        src_frame = tdcs.slip_frame(tri)
        n_rec = rec_tri.shape[0] * K
    lo, hi = (0, n_rec) if rows is None else (int(rows[0]), int(rows[1]))
    obs, tri, e = _prep(np.asarray(pts, float)[lo:hi], tri, eps)
    rec = tdcs.component_index(receiver)
    src = tdcs.component_index(source)
    out = np.zeros((hi - lo, n_src_dof, len(rec), len(src)))
    return A.interaction(obs, rec_frame[lo:hi, 2, :], rec_frame[lo:hi][:, rec, :], tri,
                         src_frame[:, src, :], mu, lam, e, out, subtract_eigenstress, order, ids, fold)
