"""Per-source records for the kernels, and the Gauss-Legendre tables.

One row of ``MHS_SRC_STRIDE`` float64 per source triangle, built from the
vertices alone (vectorised over sources):

* the centroid frame of the triangle and of its reflection in z = 0
  (src/mhs/fullspace/frame.py:67-87 applied to tri and tri * (1, 1, -1));
* the barycentric affine rows of both (shape.py:231-235);
* the fold data: straddle flag, crease centre and radius, diameter
  (src/mhs/kernels/fold.py:84-107, 344-345).

Field offsets are the ``SrcOff`` enum of ``csrc/core.hpp``.
"""
from __future__ import annotations

import functools

import numpy as np

from . import defaults

SRC_STRIDE = 80
S_V, S_EPS = 0, 9
S_E1, S_E2, S_N, S_C, S_P, S_AREA, S_L = 10, 13, 16, 19, 22, 28, 29
S_IE1, S_IE2, S_IN, S_IC = 30, 33, 36, 39
S_ABC, S_CREASE, S_CREASE_R, S_DIAM, S_STRADDLE = 42, 51, 53, 54, 55
S_IP, S_IABC, S_IAREA, S_IL, S_RC = 56, 62, 71, 72, 73

REFLECT = np.array([1.0, 1.0, -1.0])


def _frames(tri: np.ndarray):
    """(e1, e2, n, c, p (n,3,2), area, L) for a batch of triangles."""
    a = tri[:, 1] - tri[:, 0]
    b = tri[:, 2] - tri[:, 0]
    n = np.cross(a, b)
    twice = np.linalg.norm(n, axis=1)
    L = np.max(np.stack([np.linalg.norm(tri[:, 1] - tri[:, 0], axis=1),
                         np.linalg.norm(tri[:, 2] - tri[:, 1], axis=1),
                         np.linalg.norm(tri[:, 0] - tri[:, 2], axis=1)], axis=1), axis=1)
    bad = (L == 0.0) | (0.5 * twice < 1e-14 * L * L)
    if np.any(bad):
        s = int(np.argmax(bad))
        raise ValueError(f"tris[{s}] is degenerate (zero or near-zero area)")
    nhat = n / twice[:, None]
    e1 = a / np.linalg.norm(a, axis=1)[:, None]
    e2 = np.cross(nhat, e1)
    c = tri.mean(axis=1)
    d = tri - c[:, None, :]
    p = np.stack([np.einsum("svi,si->sv", d, e1), np.einsum("svi,si->sv", d, e2)], axis=2)
    return e1, e2, nhat, c, p, 0.5 * twice, L


def _affine(p: np.ndarray) -> np.ndarray:
    """Rows (A_k, B_k, C_k) of inv([1; p^T]) per source, (n, 3, 3)."""
    A = np.ones((p.shape[0], 3, 3))
    A[:, 1, :] = p[:, :, 0]
    A[:, 2, :] = p[:, :, 1]
    return np.linalg.inv(A)


def source_records(tris: np.ndarray, eps_arr: np.ndarray) -> np.ndarray:
    """(n_src, SRC_STRIDE) float64 records."""
    tri = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    n = tri.shape[0]
    rec = np.zeros((n, SRC_STRIDE))
    rec[:, S_V:S_V + 9] = tri.reshape(n, 9)
    rec[:, S_EPS] = eps_arr
    e1, e2, nh, c, p, area, L = _frames(tri)
    rec[:, S_E1:S_E1 + 3] = e1
    rec[:, S_E2:S_E2 + 3] = e2
    rec[:, S_N:S_N + 3] = nh
    rec[:, S_C:S_C + 3] = c
    rec[:, S_P:S_P + 6] = p.reshape(n, 6)
    rec[:, S_AREA] = area
    rec[:, S_L] = L
    rec[:, S_ABC:S_ABC + 9] = _affine(p).reshape(n, 9)
    ie1, ie2, inh, ic, ip, iarea, iL = _frames(tri * REFLECT)
    rec[:, S_IE1:S_IE1 + 3] = ie1
    rec[:, S_IE2:S_IE2 + 3] = ie2
    rec[:, S_IN:S_IN + 3] = inh
    rec[:, S_IC:S_IC + 3] = ic
    rec[:, S_IP:S_IP + 6] = ip.reshape(n, 6)
    rec[:, S_IABC:S_IABC + 9] = _affine(ip).reshape(n, 9)
    rec[:, S_IAREA] = iarea
    rec[:, S_IL] = iL
    # fold data (fold.py:84-107, 344-345)
    depth = -tri[:, :, 2]
    fold_d = defaults.FOLD_SOURCE_DEPTH_EPS * eps_arr
    strad = depth.min(axis=1) < fold_d
    rec[:, S_STRADDLE] = strad.astype(float)
    top = depth < fold_d[:, None]
    none = ~top.any(axis=1)
    top[none] = depth[none] == depth[none].min(axis=1, keepdims=True)
    w = top.astype(float)
    cnt = w.sum(axis=1)
    centre = np.einsum("sv,svi->si", w, tri[:, :, :2]) / cnt[:, None]
    spread = np.where(top, np.linalg.norm(tri[:, :, :2] - centre[:, None, :], axis=2), -np.inf).max(axis=1)
    top_depth = np.where(top, depth, -np.inf).max(axis=1)
    rec[:, S_CREASE:S_CREASE + 2] = centre
    rec[:, S_CREASE_R] = spread + defaults.FOLD_CREASE_MARGIN_EPS * eps_arr + top_depth
    diam = np.max(np.stack([np.linalg.norm(tri[:, i] - tri[:, j], axis=1)
                            for i in range(3) for j in range(3)], axis=1), axis=1)
    rec[:, S_DIAM] = diam
    # bounding radius about the centroid (the same for the mirrored triangle)
    rec[:, S_RC] = np.linalg.norm(tri - c[:, None, :], axis=2).max(axis=1)
    return rec


GQ_TOTAL = 2377760      # sum of n^2 for n = 1..192


@functools.lru_cache(maxsize=1)
def gauss_triangle_tables():
    """The collapsed Gauss rules on the reference triangle for orders
    1..192, flattened: order n's n^2 points at (n-1) n (2n-1) / 6 in three
    arrays x1 | x2 | w (x1 = u_i, x2 = u_j (1 - u_i), w = w_i w_j (1 - u_i),
    point g = i n + j), the same points and order as the kernels' nested loops
    (moments.py:427-498, image.py:315-377)."""
    gx, gw = gauss_tables()
    out = np.empty((3, GQ_TOTAL))
    off = 0
    for n in range(1, defaults.IMAGE_Q_GAUSS_MAX + 1):
        o1 = n * (n - 1) // 2
        u, w = gx[o1:o1 + n], gw[o1:o1 + n]
        ui = np.repeat(u, n)
        wi = np.repeat(w, n)
        uj = np.tile(u, n)
        wj = np.tile(w, n)
        out[0, off:off + n * n] = ui
        out[1, off:off + n * n] = uj * (1.0 - ui)
        out[2, off:off + n * n] = wi * wj * (1.0 - ui)
        off += n * n
    assert off == GQ_TOTAL
    return out.reshape(-1)


@functools.lru_cache(maxsize=1)
def gauss_tables():
    """(nodes, weights) on [0, 1] for orders 1..192, order n at n(n-1)/2."""
    xs, ws = [], []
    for n in range(1, defaults.IMAGE_Q_GAUSS_MAX + 1):
        x, w = np.polynomial.legendre.leggauss(n)
        xs.append(0.5 * (x + 1.0))
        ws.append(0.5 * w)
    return np.concatenate(xs), np.concatenate(ws)
