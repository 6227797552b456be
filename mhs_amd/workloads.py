"""Synthetic meshes and observer sets of the five benchmark workloads.

All seeded (``numpy.random.default_rng(20261203 + k)`` for workload k) and
wound with upward (n_z >= 0) normals; horizontal-normal (vertical) faults are
wound consistently.  Choices the workload statements leave open are fixed
here and stated in the docstrings.
"""
from __future__ import annotations

import numpy as np

SEED = 20261203


def grid_mesh(P: np.ndarray) -> np.ndarray:
    """Split every cell of a (nu+1, nv+1, 3) vertex grid into two triangles,
    wound consistently: (i,j) (i+1,j) (i+1,j+1) and (i,j) (i+1,j+1) (i,j+1)."""
    a = P[:-1, :-1].reshape(-1, 3)
    b = P[1:, :-1].reshape(-1, 3)
    c = P[1:, 1:].reshape(-1, 3)
    d = P[:-1, 1:].reshape(-1, 3)
    t1 = np.stack([a, b, c], axis=1)
    t2 = np.stack([a, c, d], axis=1)
    return np.stack([t1, t2], axis=1).reshape(-1, 3, 3)


def orient_up(tris: np.ndarray) -> np.ndarray:
    """Flip the winding of every triangle whose normal points down."""
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = n[:, 2] < 0.0
    out = tris.copy()
    out[flip] = tris[flip][:, [0, 2, 1]]
    return out


def plane(length, width, nu, nv, dip_deg, top_z, x0=0.0, y0=0.0, strike_deg=0.0):
    """Vertex grid (nu+1, nv+1, 3) of a planar fault: along strike over
    [x0 - length/2, x0 + length/2], down dip from its top edge at ``top_z``."""
    s = np.linspace(-0.5 * length, 0.5 * length, nu + 1)
    d = np.linspace(0.0, width, nv + 1)
    S, Dd = np.meshgrid(s, d, indexing="ij")
    dip = np.deg2rad(dip_deg)
    st = np.deg2rad(strike_deg)
    along = np.array([np.cos(st), np.sin(st), 0.0])
    down = np.array([-np.sin(st) * np.cos(dip), np.cos(st) * np.cos(dip), -np.sin(dip)])
    P = (np.array([x0, y0, top_z])[None, None, :] + S[..., None] * along
         + Dd[..., None] * down)
    return P


def c1_disp():
    """C1: plane 20 km (x) x 10 km down dip, dip 45 deg, top edge at z = -1,
    16 x 8 cells -> 256 triangles; 32 x 32 surface observers over
    [-30, 30]^2; eps = 0.1; Material(30, 30)."""
    P = plane(20.0, 10.0, 16, 8, 45.0, -1.0)
    tris = orient_up(grid_mesh(P))
    g = np.linspace(-30.0, 30.0, 32)
    X, Y = np.meshgrid(g, g, indexing="ij")
    obs = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    return obs, tris, 0.1


def _grf(points, ell, rng):
    """Squared-exponential Gaussian random field sample, scaled to max |g| = 1."""
    d2 = np.sum((points[:, None, :] - points[None, :, :]) ** 2, axis=2)
    C = np.exp(-0.5 * d2 / ell ** 2) + 1e-8 * np.eye(points.shape[0])
    g = np.linalg.cholesky(C) @ rng.standard_normal(points.shape[0])
    return g / np.max(np.abs(g))


def c2_stress(n_obs=4000):
    """C2: vertical base plane y = 0, x in [-20, 20], z in [-20.5, -0.5],
    50 x 20 cells -> 2000 triangles, offset normal to the plane (along y) by
    2 km g(x, z) with g a squared-exponential field (5 km) scaled to max 1.
    Per-source eps = U[0.05, 0.2] x the triangle's mean edge.  Observers
    uniform in [-30, 30]^2 x [-20, 0]."""
    rng = np.random.default_rng(SEED + 2)
    x = np.linspace(-20.0, 20.0, 51)
    z = np.linspace(-0.5, -20.5, 21)
    X, Z = np.meshgrid(x, z, indexing="ij")
    g = _grf(np.stack([X.ravel(), Z.ravel()], axis=1), 5.0, rng).reshape(X.shape)
    P = np.stack([X, 2.0 * g, Z], axis=-1)
    tris = orient_up(grid_mesh(P))
    edges = np.linalg.norm(tris - np.roll(tris, 1, axis=1), axis=2).mean(axis=1)
    eps = rng.uniform(0.05, 0.2, tris.shape[0]) * edges
    obs = np.column_stack([rng.uniform(-30, 30, n_obs), rng.uniform(-30, 30, n_obs),
                           rng.uniform(-20, 0, n_obs)])
    return obs, tris, eps


def c3_interaction(nx=200, nz=25):
    """C3: vertical strike-slip fault y = 0, x in [-50, 50], z in [-15, 0],
    200 x 25 cells (0.5 x 0.6 km) -> 10,000 triangles.  Interior vertices are
    jittered uniformly by +-5 % of the cell edge in x, y and z; boundary
    vertices stay put, so the top row touches z = 0 and the fold path runs.
    eps = 0.1, Material(30, 30)."""
    rng = np.random.default_rng(SEED + 3)
    x = np.linspace(-50.0, 50.0, nx + 1)
    z = np.linspace(0.0, -15.0, nz + 1)
    X, Z = np.meshgrid(x, z, indexing="ij")
    P = np.stack([X, np.zeros_like(X), Z], axis=-1)
    hx, hz = 100.0 / nx, 15.0 / nz
    J = rng.uniform(-0.05, 0.05, P.shape) * np.array([hx, 0.5 * (hx + hz), hz])
    J[0, :] = J[-1, :] = 0.0
    J[:, 0] = J[:, -1] = 0.0
    tris = grid_mesh(P + J)
    return tris, 0.1


def c4_network(target=6000):
    """C4: fault A 30 x 15 km dipping 60 deg, fault B 30 x 15 km dipping 80 deg
    whose plane crosses A along a line, and a 20 x 10 km splay; tops buried at
    z = -1 km (the fold is not exercised).  Cell size chosen so the total is
    about 6000 triangles.  Per-source eps = U[0.08, 0.15] km.  Returns
    (tris, groups, eps)."""
    rng = np.random.default_rng(SEED + 4)
    h = np.sqrt((2 * 450.0 + 200.0) / target * 2.0)
    meshes, groups = [], []
    specs = [(30.0, 15.0, 60.0, 0.0, 0.0, 0.0),
             (30.0, 15.0, 80.0, 0.0, 4.0, 25.0),
             (20.0, 10.0, 45.0, 12.0, -6.0, -30.0)]
    for gid, (L, W, dip, x0, y0, strike) in enumerate(specs):
        nu = max(1, int(round(L / h)))
        nv = max(1, int(round(W / h)))
        P = plane(L, W, nu, nv, dip, -1.0, x0, y0, strike)
        t = orient_up(grid_mesh(P))
        meshes.append(t)
        groups.append(np.full(t.shape[0], gid))
    tris = np.concatenate(meshes)
    return tris, np.concatenate(groups), rng.uniform(0.08, 0.15, tris.shape[0])


def c5_request(rng, kind_u=None):
    """One request of the served stream: a random-perturbed planar fault with
    n ~ log-uniform[100, 8000] triangles, dip U[20, 90] deg, top depth
    U[0, 5] km; kind 40 % disp (32 x 32 surface grid), 30 % traction (2000
    observers, random unit normals), 30 % interaction; order U{0, 1, 2}."""
    n = int(np.exp(rng.uniform(np.log(100), np.log(8000))))
    dip = rng.uniform(20.0, 90.0)
    top = -rng.uniform(0.0, 5.0)
    aspect = 2.0
    nv = max(1, int(round(np.sqrt(n / (2 * aspect)))))
    nu = max(1, int(round(n / (2 * nv))))
    L = 2.0 * nu
    W = 2.0 * nv * 0.5
    P = plane(L, W, nu, nv, dip, top)
    hx = L / nu
    J = rng.uniform(-0.05, 0.05, P.shape) * hx
    J[:, 0, 2] = 0.0
    tris = orient_up(grid_mesh(P + J))
    tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
    u = rng.uniform() if kind_u is None else kind_u
    kind = "disp" if u < 0.4 else ("traction" if u < 0.7 else "interaction")
    order = int(rng.integers(0, 3))
    return {"tris": tris, "kind": kind, "order": order, "eps": 0.1}
