"""Shared-node table for continuous P2 slip (the 'continuous' layout).

Mirrors src/mhs/shared.py:81-279: vertices welded per group at
tol_over_h x the shortest edge, one node per edge, ids (n_tri, 6) in the order
v1 v2 v3 m12 m23 m31, area-weighted upward node normals, boundary, quadratic
integration weights; refuses non-manifold, inconsistent winding and soups.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import tdcs

CONTINUOUS_ORDER = 2
LAYOUTS = ("discontinuous", "continuous")
WELD_TOL_OVER_H = 1.0e-9


@dataclasses.dataclass(frozen=True)
class SharedNodes:
    coords: np.ndarray
    ids: np.ndarray
    normals: np.ndarray
    boundary: np.ndarray
    weights: np.ndarray
    group: np.ndarray
    n_vertices: int
    tol: float

    @property
    def n(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_tri(self) -> int:
        return int(self.ids.shape[0])

    @property
    def n_edges(self) -> int:
        return self.n - self.n_vertices

    @property
    def frames(self) -> np.ndarray:
        return tdcs.frame_from_normals(self.normals)

    @property
    def trace(self) -> np.ndarray:
        return self.boundary & (self.coords[:, 2] >= -self.tol)

    @property
    def front(self) -> np.ndarray:
        return self.boundary & ~self.trace

    def expand(self, values) -> np.ndarray:
        v = np.asarray(values)
        if v.shape[0] != self.n:
            raise ValueError(f"expected {self.n} nodal values, got {v.shape[0]}")
        return v[self.ids].reshape((self.n_tri * 6,) + v.shape[1:])

    def integrate(self, values) -> np.ndarray:
        v = np.asarray(values, float)
        if v.shape[0] != self.n:
            raise ValueError(f"expected {self.n} nodal values, got {v.shape[0]}")
        return np.tensordot(self.weights, v, axes=(0, 0))

    def element_mean(self, values) -> np.ndarray:
        v = np.asarray(values, float)
        return v[self.ids[:, 3:]].mean(axis=1)


def _edges(idx):
    local = np.stack([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]], axis=1)
    flat = np.sort(local.reshape(-1, 2), axis=1)
    edges, inv, counts = np.unique(flat, axis=0, return_inverse=True, return_counts=True)
    return edges, inv.reshape(-1, 3), counts


def shared_nodes(tris, *, groups=None, tol_over_h=WELD_TOL_OVER_H) -> SharedNodes:
    """Build the table (shared.py:175-273)."""
    tri = np.ascontiguousarray(np.asarray(tris, dtype=np.float64))
    if tri.ndim != 3 or tri.shape[1:] != (3, 3) or tri.shape[0] == 0:
        raise ValueError(f"tris must have shape (n_tri, 3, 3), got {tri.shape}")
    if not np.all(np.isfinite(tri)):
        raise ValueError("tris must be finite")
    n_tri = tri.shape[0]
    grp = (np.zeros(n_tri, dtype=np.int64) if groups is None
           else np.asarray(groups).reshape(-1).astype(np.int64))
    if grp.shape[0] != n_tri:
        raise ValueError(f"groups has {grp.shape[0]} entries for {n_tri} triangles")
    h_min = float(np.linalg.norm(tri - np.roll(tri, 1, axis=1), axis=2).min())
    if not h_min > 0.0:
        raise ValueError("a triangle has a zero-length edge")
    tol = float(tol_over_h) * h_min
    flat = tri.reshape(-1, 3)
    keys = np.concatenate([np.round(flat / tol).astype(np.int64),
                           np.repeat(grp, 3)[:, None]], axis=1)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    n_v = uniq.shape[0]
    verts = np.zeros((n_v, 3))
    np.add.at(verts, inv, flat)
    verts /= np.bincount(inv, minlength=n_v)[:, None]
    idx = inv.reshape(-1, 3)
    if (np.any(idx[:, 0] == idx[:, 1]) or np.any(idx[:, 1] == idx[:, 2])
            or np.any(idx[:, 2] == idx[:, 0])):
        raise ValueError(f"a triangle has two vertices within the weld tolerance "
                         f"{tol:.3e}: degenerate, or tol_over_h too loose")
    edges, edge_of, counts = _edges(idx)
    if np.any(counts > 2):
        bad = int(np.argmax(counts))
        raise ValueError(f"non-manifold: edge {edges[bad].tolist()} is shared by "
                         f"{int(counts[bad])} triangles")
    if n_tri > 1 and not np.any(counts == 2):
        raise ValueError(f"nothing is shared: {n_tri} triangles and no common edge; "
                         f"a soup wants the discontinuous layout")
    local = np.stack([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]], axis=1).reshape(-1, 2)
    forward = local[:, 0] < local[:, 1]
    n_forward = np.bincount(edge_of.reshape(-1), weights=forward, minlength=edges.shape[0])
    bad_wind = (counts == 2) & (n_forward != 1)
    if np.any(bad_wind):
        bad = int(np.argmax(bad_wind))
        raise ValueError(f"inconsistent winding: the two triangles on edge "
                         f"{edges[bad].tolist()} traverse it the same way")
    n_e = edges.shape[0]
    mids = 0.5 * (verts[edges[:, 0]] + verts[edges[:, 1]])
    coords = np.concatenate([verts, mids])
    ids = np.concatenate([idx, n_v + edge_of], axis=1)
    boundary = np.zeros(n_v + n_e, bool)
    rim = counts == 1
    boundary[n_v + np.flatnonzero(rim)] = True
    boundary[edges[rim].reshape(-1)] = True
    nrm_el = tdcs.unit_normals(tri)
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    normals = np.zeros((n_v + n_e, 3))
    np.add.at(normals, ids.reshape(-1), np.repeat(area[:, None] * nrm_el, 6, axis=0))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    weights = np.zeros(n_v + n_e)
    np.add.at(weights, ids[:, 3:].reshape(-1), np.repeat(area / 3.0, 3))
    group = np.zeros(n_v + n_e, dtype=np.int64)
    group[ids.reshape(-1)] = np.repeat(grp, 6)
    return SharedNodes(coords=coords, ids=ids, normals=normals, boundary=boundary,
                       weights=weights, group=group, n_vertices=int(n_v), tol=tol)


def check_layout(layout: str) -> str:
    if layout not in LAYOUTS:
        raise ValueError(f"layout must be one of {LAYOUTS}, got {layout!r}")
    return layout
