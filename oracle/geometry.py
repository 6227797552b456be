"""Triangle frame, Lagrange shape coefficients and the Gauss rules (oracle).

Restates src/mhs/fullspace/frame.py:67-87 (local frame), shape.py:164-271
(lattice, nodes, shape coefficients) and moments.py:401-424 (the collapsed
triangle rule).
"""
from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np

DEGENERATE_AREA_REL = 1e-14      # fullspace/defaults.py:37


@dataclass(frozen=True)
class Frame:
    v: np.ndarray          # (3, 3)
    e1: np.ndarray
    e2: np.ndarray
    nhat: np.ndarray
    centroid: np.ndarray
    p: np.ndarray          # (3, 2) vertices in the centroid frame
    area: float
    L: float               # longest edge

    def to_plane(self, pts):
        """(z, X): signed height and in-plane coords (frame.py:36-41)."""
        d = np.asarray(pts, float).reshape(-1, 3) - self.centroid
        return d @ self.nhat, np.stack([d @ self.e1, d @ self.e2], axis=1)


def frame_of(tri) -> Frame:
    """Centroid frame of one triangle (frame.py:67-87)."""
    tri = np.asarray(tri, float).reshape(3, 3)
    a = tri[1] - tri[0]
    b = tri[2] - tri[0]
    n = np.cross(a, b)
    twice_area = np.linalg.norm(n)
    L = float(max(np.linalg.norm(tri[1] - tri[0]), np.linalg.norm(tri[2] - tri[1]),
                  np.linalg.norm(tri[0] - tri[2])))
    if L == 0.0 or 0.5 * twice_area < DEGENERATE_AREA_REL * L * L:
        raise ValueError("degenerate triangle (zero or near-zero area)")
    nhat = n / twice_area
    e1 = a / np.linalg.norm(a)
    e2 = np.cross(nhat, e1)
    c = tri.mean(axis=0)
    d = tri - c
    return Frame(tri, e1, e2, nhat, c, np.stack([d @ e1, d @ e2], axis=1),
                 0.5 * twice_area, L)


def n_nodes(p: int) -> int:
    return (p + 1) * (p + 2) // 2


def lattice(p: int) -> np.ndarray:
    """Node multi-indices in the order v1 v2 v3 m12 m23 m31 (shape.py:179-202);
    only p <= 2 is needed by the matrix entry points."""
    if p == 0:
        return np.array([[0, 0, 0]])
    if p == 1:
        return np.eye(3, dtype=int)
    if p == 2:
        return np.array([[2, 0, 0], [0, 2, 0], [0, 0, 2],
                         [1, 1, 0], [0, 1, 1], [1, 0, 1]])
    raise ValueError("order must be 0, 1 or 2")


def _poly_mul(P, Q, deg):
    """Truncated product of (..., deg+1, deg+1) coefficient arrays
    (shape.py:218-228)."""
    out = np.zeros_like(P)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            if not np.any(P[..., a, b]):
                continue
            for c in range(deg + 1 - a):
                for d in range(deg + 1 - a - b - c):
                    out[..., a + c, b + d] += P[..., a, b] * Q[..., c, d]
    return out


def bary_affine(fr: Frame) -> np.ndarray:
    """Rows (A_k, B_k, C_k): lam_k = A + B eta1 + C eta2 (shape.py:231-235)."""
    return np.linalg.inv(np.vstack([np.ones(3), fr.p.T]))


def shape_coefficients(fr: Frame, p: int, X) -> np.ndarray:
    """(N, K, p+1, p+1) monomial coefficients of N_k in xi = eta - X
    (shape.py:238-271)."""
    X = np.asarray(X, float).reshape(-1, 2)
    N = X.shape[0]
    K = n_nodes(p)
    D = p + 1
    if p == 0:
        c = np.zeros((N, 1, 1, 1))
        c[:, 0, 0, 0] = 1.0
        return c
    abc = bary_affine(fr)
    lam = np.zeros((N, 3, D, D))
    lam[:, :, 0, 0] = abc[:, 0][None] + X[:, 0:1] * abc[:, 1][None] + X[:, 1:2] * abc[:, 2][None]
    lam[:, :, 1, 0] = abc[:, 1][None]
    lam[:, :, 0, 1] = abc[:, 2][None]
    c = np.zeros((N, K, D, D))
    for k, alpha in enumerate(lattice(p)):
        poly = np.zeros((N, D, D))
        poly[:, 0, 0] = 1.0
        for i in range(3):
            for j in range(int(alpha[i])):
                fac = p * lam[:, i] / (j + 1)
                fac[:, 0, 0] -= j / (j + 1)
                poly = _poly_mul(poly, fac, p)
        c[:, k] = poly
    return c


def nodal_at(fr: Frame, p: int, eta) -> np.ndarray:
    """N_k at centroid-frame points eta (Q, 2) -> (Q, K), the way the far-field
    producer evaluates them (moments.py:441-446, image.py:331-338)."""
    c0 = shape_coefficients(fr, p, np.zeros((1, 2)))[0]
    K, D1, _ = c0.shape
    out = np.zeros((eta.shape[0], K))
    for a in range(D1):
        for b in range(D1 - a):
            out += c0[:, a, b][None, :] * (eta[:, 0] ** a * eta[:, 1] ** b)[:, None]
    return out


@functools.lru_cache(maxsize=None)
def gauss_legendre01(n: int):
    """Gauss-Legendre on [0, 1] (fold.py:238-240)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


@functools.lru_cache(maxsize=None)
def gauss_triangle(n: int):
    """Collapsed product rule on the unit triangle, sum(w) = 1/2
    (moments.py:401-424)."""
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1.0)
    w = 0.5 * w
    gi, gj = np.meshgrid(g, g, indexing="ij")
    wi, wj = np.meshgrid(w, w, indexing="ij")
    return gi.ravel(), (gj * (1.0 - gi)).ravel(), (wi * wj * (1.0 - gi)).ravel()


def unit_normal(tri) -> np.ndarray:
    """Vertex-order normal (assemble.py:369-382)."""
    v = np.asarray(tri, float)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    return n / np.linalg.norm(n)
