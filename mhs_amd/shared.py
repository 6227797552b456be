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


def _coerce(tris):
    tri = np.ascontiguousarray(np.asarray(tris, dtype=np.float64))
    if tri.ndim != 3 or tri.shape[1:] != (3, 3) or tri.shape[0] == 0:
        raise ValueError(f"tris must have shape (n_tri, 3, 3), got {tri.shape}")
    return tri


def _groups(groups, n_tri):
    grp = (np.zeros(n_tri, dtype=np.int64) if groups is None
           else np.asarray(groups).reshape(-1).astype(np.int64))
    if grp.shape[0] != n_tri:
        raise ValueError(f"groups has {grp.shape[0]} entries for {n_tri} triangles")
    return np.ascontiguousarray(grp)


def _prepare(tris, groups, tol_over_h):
    """Validated float64 triangles, int64 groups and the weld tolerance."""
    tri = _coerce(tris)
    if not np.all(np.isfinite(tri)):
        raise ValueError("tris must be finite")
    grp = _groups(groups, tri.shape[0])
    h_min = float(np.linalg.norm(tri - np.roll(tri, 1, axis=1), axis=2).min())
    if not h_min > 0.0:
        raise ValueError("a triangle has a zero-length edge")
    return tri, grp, float(tol_over_h) * h_min


def _refuse_degenerate(idx, tol):
    if (np.any(idx[:, 0] == idx[:, 1]) or np.any(idx[:, 1] == idx[:, 2])
            or np.any(idx[:, 2] == idx[:, 0])):
        raise ValueError(f"a triangle has two vertices within the weld tolerance "
                         f"{tol:.3e}: degenerate, or tol_over_h too loose")


def _refuse_edges(edges, counts, n_forward, n_tri):
    """Non-manifold edges, a soup, inconsistent winding (shared.py:213-237)."""
    if np.any(counts > 2):
        bad = int(np.argmax(counts))
        raise ValueError(f"non-manifold: edge {edges[bad].tolist()} is shared by "
                         f"{int(counts[bad])} triangles")
    if n_tri > 1 and not np.any(counts == 2):
        raise ValueError(f"nothing is shared: {n_tri} triangles and no common edge; "
                         f"a soup wants the discontinuous layout")
    bad_wind = (counts == 2) & (n_forward != 1)
    if np.any(bad_wind):
        bad = int(np.argmax(bad_wind))
        raise ValueError(f"inconsistent winding: the two triangles on edge "
                         f"{edges[bad].tolist()} traverse it the same way")


def _area(tri):
    return 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)


def shared_nodes(tris, *, groups=None, tol_over_h=WELD_TOL_OVER_H) -> SharedNodes:
    """Build the table (shared.py:175-273)."""
    tri, grp, tol = _prepare(tris, groups, tol_over_h)
    n_tri = tri.shape[0]
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
    _refuse_degenerate(idx, tol)
    edges, edge_of, counts = _edges(idx)
    local = np.stack([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]], axis=1).reshape(-1, 2)
    forward = local[:, 0] < local[:, 1]
    n_forward = np.bincount(edge_of.reshape(-1), weights=forward, minlength=edges.shape[0])
    _refuse_edges(edges, counts, n_forward, n_tri)
    n_e = edges.shape[0]
    mids = 0.5 * (verts[edges[:, 0]] + verts[edges[:, 1]])
    coords = np.concatenate([verts, mids])
    ids = np.concatenate([idx, n_v + edge_of], axis=1)
    boundary = np.zeros(n_v + n_e, bool)
    rim = counts == 1
    boundary[n_v + np.flatnonzero(rim)] = True
    boundary[edges[rim].reshape(-1)] = True
    nrm_el = tdcs.unit_normals(tri)
    area = _area(tri)
    normals = np.zeros((n_v + n_e, 3))
    np.add.at(normals, ids.reshape(-1), np.repeat(area[:, None] * nrm_el, 6, axis=0))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    weights = np.zeros(n_v + n_e)
    np.add.at(weights, ids[:, 3:].reshape(-1), np.repeat(area / 3.0, 3))
    group = np.zeros(n_v + n_e, dtype=np.int64)
    group[ids.reshape(-1)] = np.repeat(grp, 6)
    return SharedNodes(coords=coords, ids=ids, normals=normals, boundary=boundary,
                       weights=weights, group=group, n_vertices=int(n_v), tol=tol)


def shared_nodes_device(tris, *, groups=None, tol_over_h=WELD_TOL_OVER_H) -> SharedNodes:
    """:func:`shared_nodes` built on the GPU, for fault networks whose host
    ``np.unique`` passes dominate: the same arguments, table, refusals and
    messages.  The per-triangle geometry (``mhs_tri_geometry``: shortest
    edge, element normals and areas), the weld, the edge table, the per-node
    sums and the refusal checks (``mhs_shared_nodes``) run on the device;
    the edge table comes back to the host only to word a refusal.  Node ids,
    the boundary and the groups equal the host table's, and coordinates,
    normals and weights are summed in the host's order with the same IEEE
    operations."""
    import ctypes as C

    from . import _lib, runtime
    tri = _coerce(tris)
    n_tri = tri.shape[0]
    dev = runtime.device()
    torch, lib = dev.torch, dev.lib
    n3 = 3 * n_tri
    work_bytes = int(lib.mhs_weld_workspace_bytes(n_tri))
    if work_bytes < 0:
        raise ValueError(f"shared_nodes_device: {n_tri} triangles is out of range (3 n_tri < 2^31)")
    stream = torch.cuda.current_stream(dev.dev)
    h_st = C.c_void_p(stream.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    f64, i64 = torch.float64, torch.int64

    d_tri = torch.from_numpy(tri).to(dev.dev)
    area_normals = torch.empty((n_tri, 3), dtype=f64, device=dev.dev)
    area3 = torch.empty(n_tri, dtype=f64, device=dev.dev)
    stats_dev = torch.empty(4, dtype=i64, device=dev.dev)
    stats = (C.c_double * 4)()
    _lib.check(lib.mhs_tri_geometry(ptr(d_tri), n_tri, ptr(area_normals), ptr(area3), ptr(stats_dev), stats,
                                    h_st), lib)
    h_min, coord_max, n_nonfinite, n_collinear = stats[0], stats[1], int(stats[2]), int(stats[3])
    if n_nonfinite:
        raise ValueError("tris must be finite")
    grp = _groups(groups, n_tri)
    if not h_min > 0.0:
        raise ValueError("a triangle has a zero-length edge")
    tol = float(tol_over_h) * h_min
    if not coord_max / tol < 2.0 ** 62:
        raise ValueError(f"coordinates too large for the weld tolerance {tol:.3e}: "
                         f"the weld keys would overflow int64")

    d_grp = torch.from_numpy(grp).to(dev.dev)
    work = torch.empty(work_bytes, dtype=torch.uint8, device=dev.dev)
    ids = torch.empty((n_tri, 6), dtype=i64, device=dev.dev)
    coords = torch.empty((2 * n3, 3), dtype=f64, device=dev.dev)
    normals = torch.empty((2 * n3, 3), dtype=f64, device=dev.dev)
    weights = torch.empty(2 * n3, dtype=f64, device=dev.dev)
    group = torch.empty(2 * n3, dtype=i64, device=dev.dev)
    boundary = torch.empty(2 * n3, dtype=torch.uint8, device=dev.dev)
    edges = torch.empty((n3, 2), dtype=i64, device=dev.dev)
    counts = torch.empty(n3, dtype=torch.int32, device=dev.dev)
    n_forward = torch.empty(n3, dtype=torch.int32, device=dev.dev)
    n_nodes = (C.c_int64 * 2)()
    refusals = C.c_int(0)
    _lib.check(lib.mhs_shared_nodes(ptr(d_tri), ptr(d_grp), n_tri, tol, ptr(area_normals), ptr(area3), ptr(work),
                                    work_bytes, ptr(ids), ptr(coords), ptr(normals), ptr(weights),
                                    ptr(group), ptr(boundary), ptr(edges), ptr(counts), ptr(n_forward),
                                    n_nodes, C.byref(refusals), h_st), lib)
    n_v, n_e = int(n_nodes[0]), int(n_nodes[1])
    n = n_v + n_e
    bits = refusals.value
    h_ids = _host_i64(runtime, ids)
    if bits & _lib.REFUSE_DEGENERATE:
        _refuse_degenerate(h_ids[:, :3], tol)
        raise RuntimeError("mhs_shared_nodes flagged a degenerate triangle the id table does not show")
    if (bits & (_lib.REFUSE_NON_MANIFOLD | _lib.REFUSE_WINDING)
            or (n_tri > 1 and not bits & _lib.REFUSE_SHARED)):
        _refuse_edges(edges[:n_e].cpu().numpy(), counts[:n_e].cpu().numpy(), n_forward[:n_e].cpu().numpy(), n_tri)
        raise RuntimeError("mhs_shared_nodes flagged an edge refusal the edge table does not show")
    if n_collinear:
        tdcs.unit_normals(tri)          # raises the host's message for the first collinear triangle
    return SharedNodes(coords=runtime.to_host(coords[:n]), ids=h_ids, normals=runtime.to_host(normals[:n]),
                       boundary=boundary[:n].cpu().numpy().astype(bool), weights=runtime.to_host(weights[:n]),
                       group=_host_i64(runtime, group[:n]), n_vertices=n_v, tol=tol)


def _host_i64(runtime, t):
    """An int64 device tensor on the host through the pinned staged copy
    (bit for bit: the words travel as float64 and are only copied)."""
    out = np.empty(tuple(t.shape), dtype=np.int64)
    runtime.to_host(t.view(runtime.device().torch.float64), out.view(np.float64))
    return out


def check_layout(layout: str) -> str:
    if layout not in LAYOUTS:
        raise ValueError(f"layout must be one of {LAYOUTS}, got {layout!r}")
    return layout
