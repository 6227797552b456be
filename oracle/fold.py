"""Folded image of a source whose smeared slip reaches above z = 0.

Restates src/mhs/kernels/fold.py:75-380 with the constants of
src/mhs/defaults.py:108-148: sources with their shallowest vertex within
3 eps of the surface are folded at receivers shallower than 10 eps and within
60 eps (horizontally) of the source's crease; there the image is the mirror
element's full-space field (slip columns signed (-1, -1, +1)) plus a
Boussinesq surface integral of the pair's normal traction on a receiver-
centred polar grid.
"""
from __future__ import annotations

import numpy as np

from .fullspace import blocks
from .geometry import gauss_legendre01, n_nodes, unit_normal

REFLECT = np.array([1.0, 1.0, -1.0])
MIRROR_SLIP_SIGN = np.array([-1.0, -1.0, 1.0])
ZHAT = np.array([0.0, 0.0, 1.0])
FOLD_SOURCE_DEPTH_EPS = 3.0
FOLD_BAND_DEPTH_EPS = 10.0
FOLD_REACH_EPS = 60.0
FOLD_N_RADIAL = 6
FOLD_N_ANGULAR = 4
FOLD_CREASE_MARGIN_EPS = 5.0
FOLD_FINE_SPACING_EPS = 0.25
FOLD_MEDIUM_SPACING_EPS = 2.0
FOLD_N_COARSE = 24
FOLD_FINE_SUBPANEL_EPS = 6.0
FOLD_MEDIUM_SUBPANEL_EPS = 16.0
FOLD_R1_MARGIN_DIAM = 8.0


def straddles(tri, eps) -> bool:
    """fold.py:84-87."""
    return bool(float(-np.asarray(tri, float)[:, 2].max()) < FOLD_SOURCE_DEPTH_EPS * float(eps))


def crease(tri, eps):
    """(centre, radius) of the surface disk of sharp features (fold.py:95-107)."""
    tri = np.asarray(tri, float)
    depth = -tri[:, 2]
    top = tri[depth < FOLD_SOURCE_DEPTH_EPS * eps]
    if top.shape[0] == 0:
        top = tri[depth == depth.min()]
    centre = top[:, :2].mean(axis=0)
    spread = float(np.linalg.norm(top[:, :2] - centre, axis=1).max())
    return centre, spread + FOLD_CREASE_MARGIN_EPS * eps + float((-top[:, 2]).max())


def band(obs, tri, eps) -> np.ndarray:
    """Receivers taking the folded image (fold.py:110-118)."""
    obs = np.asarray(obs, float)
    centre, _ = crease(tri, eps)
    reach = np.hypot(obs[:, 0] - centre[0], obs[:, 1] - centre[1])
    return np.flatnonzero((-obs[:, 2] < FOLD_BAND_DEPTH_EPS * eps) & (reach < FOLD_REACH_EPS * eps))


def point_stress(dx, dy, zeta, nu):
    """Boussinesq unit compressive point load, stress (fold.py:132-166)."""
    zeta = np.broadcast_to(np.asarray(zeta, float), dx.shape)
    r2 = dx * dx + dy * dy
    rho = np.sqrt(r2 + zeta * zeta)
    rho3 = rho ** 3
    rho5 = rho3 * rho * rho
    a = 1.0 / (rho * (rho + zeta))
    c = 1.0 - 2.0 * nu
    s_th = c * (-a + zeta / rho3)
    s_rth = 2.0 * c * a - 3.0 * zeta * r2 / rho5 - c * zeta / rho3
    r = np.sqrt(r2)
    with np.errstate(invalid="ignore", divide="ignore"):
        ex = np.where(r > 0, dx / r, 0.0)
        ey = np.where(r > 0, dy / r, 0.0)
    pref = 1.0 / (2.0 * np.pi)
    S = np.empty(dx.shape + (3, 3))
    S[..., 0, 0] = pref * (s_th + s_rth * ex * ex)
    S[..., 1, 1] = pref * (s_th + s_rth * ey * ey)
    S[..., 0, 1] = S[..., 1, 0] = pref * s_rth * ex * ey
    t_rz = pref * 3.0 * r * zeta * zeta / rho5
    S[..., 0, 2] = S[..., 2, 0] = t_rz * ex
    S[..., 1, 2] = S[..., 2, 1] = t_rz * ey
    S[..., 2, 2] = -pref * 3.0 * zeta ** 3 / rho5
    return S


def point_disp(dx, dy, zeta, mu, nu):
    """Boussinesq point load, displacement (fold.py:169-182)."""
    zeta = np.broadcast_to(np.asarray(zeta, float), dx.shape)
    rho = np.sqrt(dx * dx + dy * dy + zeta * zeta)
    pref = 1.0 / (4.0 * np.pi * mu)
    ur = pref * (zeta / rho ** 3 - (1.0 - 2.0 * nu) / (rho * (rho + zeta)))
    u = np.empty(dx.shape + (3,))
    u[..., 0] = ur * dx
    u[..., 1] = ur * dy
    u[..., 2] = -pref * (2.0 * (1.0 - nu) / rho + zeta * zeta / rho ** 3)
    return u


def disk_stress(a, zeta, nu):
    """Love's on-axis uniform disk load, stress (fold.py:185-198)."""
    rho = np.sqrt(a * a + zeta * zeta)
    szz = -(1.0 - zeta ** 3 / rho ** 3)
    srr = -0.5 * ((1.0 + 2.0 * nu) - 2.0 * (1.0 + nu) * zeta / rho + zeta ** 3 / rho ** 3)
    return np.diag([srr, srr, szz])


def disk_disp(a, zeta, mu, nu):
    """Love's on-axis disk load, displacement (fold.py:201-210)."""
    rho = np.sqrt(a * a + zeta * zeta)
    return np.array([0.0, 0.0, -(0.5 / mu) * (2.0 * (1.0 - nu) * (rho - zeta) + zeta - zeta * zeta / rho)])


def surface_load(pts_h, tri, mu, lam, eps, order):
    """2 (H - E C)_zz on z = 0 per node and slip: (n, K, 3) (fold.py:214-228)."""
    pts_h = np.asarray(pts_h, float)
    obs = np.column_stack([pts_h, np.zeros(pts_h.shape[0])])
    blk = blocks(obs, tri, mu, lam, eps, ("H", "E"), order)
    nrm = unit_normal(tri)
    e_zz = lam * nrm + 2.0 * mu * nrm[2] * ZHAT
    return 2.0 * (blk["H"][:, :, 2, 2, :] - blk["E"][:, :, None] * e_zz[None, None, :])


def _composite(breaks, n):
    """fold.py:243-248."""
    u, wu = gauss_legendre01(n)
    a = np.asarray(breaks[:-1], float)
    b = np.asarray(breaks[1:], float)
    return ((a[:, None] + (b - a)[:, None] * u[None, :]).ravel(),
            ((b - a)[:, None] * wu[None, :]).ravel())


def _halfwidth(ra, rb, rho_s, radius):
    """fold.py:251-258."""
    if rb < rho_s - radius or ra > rho_s + radius:
        return 0.0
    if rho_s <= radius:
        return np.pi
    return float(np.arcsin(min(1.0, radius / rho_s)))


def _angular(theta_c, w_fine, w_med, rb, eps):
    """Composite angular rule (fold.py:261-287)."""
    n = FOLD_N_ANGULAR
    if w_fine >= np.pi:
        m = max(8, int(np.ceil(2 * np.pi * rb / (n * FOLD_FINE_SPACING_EPS * eps))))
        return _composite(np.linspace(0, 2 * np.pi, m + 1), n)
    if w_med >= np.pi:
        m = max(8, int(np.ceil(2 * np.pi * rb / (n * FOLD_MEDIUM_SPACING_EPS * eps))))
        return _composite(np.linspace(0, 2 * np.pi, m + 1), n)
    pieces = []
    if w_fine > 0:
        m = max(1, int(np.ceil(2 * w_fine * rb / (n * FOLD_FINE_SPACING_EPS * eps))))
        pieces.append(np.linspace(theta_c - w_fine, theta_c + w_fine, m + 1))
    if w_med > w_fine:
        m = max(1, int(np.ceil((w_med - w_fine) * rb / (n * FOLD_MEDIUM_SPACING_EPS * eps))))
        pieces.insert(0, np.linspace(theta_c - w_med, theta_c - w_fine, m + 1))
        pieces.append(np.linspace(theta_c + w_fine, theta_c + w_med, m + 1))
    start = theta_c + w_med if w_med > 0 else 0.0
    pieces.append(np.linspace(start, start + 2 * np.pi - 2 * w_med, max(1, FOLD_N_COARSE // n) + 1))
    nodes, weights = [], []
    for br in pieces:
        if br.size >= 2 and br[-1] > br[0]:
            t, w = _composite(br, n)
            nodes.append(t)
            weights.append(w)
    return np.concatenate(nodes), np.concatenate(weights)


def polar_breaks(rho_s, zeta, r_f, eps, diam):
    """Radial panel breaks and R1 (fold.py:296-318)."""
    r1_target = rho_s + 4 * r_f + FOLD_R1_MARGIN_DIAM * diam
    geo = [0.0]
    if 0 < zeta < 0.5 * eps:
        geo += [z for z in (0.5 * zeta, zeta, 2 * zeta, 4 * zeta) if z < 0.5 * eps]
    r = 0.5 * eps
    while True:
        geo.append(r)
        if r >= r1_target:
            break
        r *= 2.0
    R1 = geo[-1]
    lo_m, hi_m = rho_s - 4 * r_f, rho_s + 4 * r_f
    lo_f, hi_f = rho_s - r_f, rho_s + r_f
    brk = [g for g in geo if g < lo_m or g > hi_m]
    ann = []
    for a0, b0, width in ((max(0.0, lo_m), max(0.0, lo_f), FOLD_MEDIUM_SUBPANEL_EPS),
                          (max(0.0, lo_f), min(R1, hi_f), FOLD_FINE_SUBPANEL_EPS),
                          (min(R1, hi_f), min(R1, hi_m), FOLD_MEDIUM_SUBPANEL_EPS)):
        if b0 > a0:
            ann += list(np.linspace(a0, b0, max(1, int(np.ceil((b0 - a0) / (width * eps)))) + 1))
    brk = np.unique(np.concatenate([np.asarray(brk, float), np.asarray(ann, float)]))
    brk = brk[brk <= R1 + 1e-12 * R1]
    return brk, R1


def polar_grid(xh, zeta, centre, r_f, eps, diam):
    """Receiver-centred surface quadrature (fold.py:290-330)."""
    xh = np.asarray(xh, float)
    centre = np.asarray(centre, float)
    rho_s = float(np.hypot(*(centre - xh)))
    theta_c = float(np.arctan2(centre[1] - xh[1], centre[0] - xh[0]))
    brk, R1 = polar_breaks(rho_s, zeta, r_f, eps, diam)
    pts, wts = [], []
    for ra, rb in zip(brk[:-1], brk[1:]):
        if rb - ra < 1e-9 * eps:
            continue
        rq, wr = _composite([ra, rb], FOLD_N_RADIAL)
        tq, wt = _angular(theta_c, _halfwidth(ra, rb, rho_s, r_f),
                          _halfwidth(ra, rb, rho_s, 4 * r_f), rb, eps)
        R, T = np.meshgrid(rq, tq, indexing="ij")
        pts.append(np.stack([xh[0] + R * np.cos(T), xh[1] + R * np.sin(T)], axis=-1).reshape(-1, 2))
        wts.append((wr[:, None] * wt[None, :] * R).ravel())
    return np.concatenate(pts), np.concatenate(wts), R1


def boussinesq(obs, tri, mu, lam, eps, order, want=("H",)):
    """Boussinesq correction of one source at every obs (fold.py:334-367)."""
    obs = np.asarray(obs, float)
    tri = np.asarray(tri, float)
    nu = 0.5 * lam / (lam + mu)
    centre, r_f = crease(tri, eps)
    diam = max(float(np.linalg.norm(tri[i] - tri[j])) for i in range(3) for j in range(3))
    K = n_nodes(order)
    s_h = surface_load(obs[:, :2], tri, mu, lam, eps, order)
    out = {}
    if "U" in want:
        out["U"] = np.empty((obs.shape[0], K, 3, 3))
    if "H" in want:
        out["H"] = np.empty((obs.shape[0], K, 3, 3, 3))
    for o in range(obs.shape[0]):
        xh = obs[o, :2]
        zeta = max(0.0, float(-obs[o, 2]))
        pts, wts, R1 = polar_grid(xh, zeta, centre, r_f, eps, diam)
        s = surface_load(pts, tri, mu, lam, eps, order) - s_h[o][None]
        dx, dy = xh[0] - pts[:, 0], xh[1] - pts[:, 1]
        if "H" in want:
            SB = point_stress(dx, dy, zeta, nu)
            out["H"][o] = (np.einsum("n,nab,nkm->kabm", wts, SB, s)
                           + s_h[o][:, None, None, :] * disk_stress(R1, zeta, nu)[None, :, :, None])
        if "U" in want:
            UB = point_disp(dx, dy, zeta, mu, nu)
            out["U"][o] = (np.einsum("n,ni,nkm->kim", wts, UB, s)
                           + s_h[o][:, None, :] * disk_disp(R1, zeta, mu, nu)[None, :, None])
    return out


def image(obs, tri, mu, lam, eps, order, want=("H",)):
    """Mirror element full space (signed slip) + Boussinesq (fold.py:370-380)."""
    blk = blocks(obs, np.asarray(tri, float) * REFLECT, mu, lam, eps, tuple(want), order)
    corr = boussinesq(obs, tri, mu, lam, eps, order, want)
    return {k: blk[k] * MIRROR_SLIP_SIGN + corr[k] for k in want}
