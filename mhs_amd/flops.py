"""Algorithmic work of an assembly: the branch histogram and FLOP model.

Every pair takes branches the reference's rules fix exactly, so the point
counts are computable on the host without running the kernels:

* direct term: closed form if sqrt(|x - c|^2 + eps^2) <= 10 L, else Gauss
  12^2 (<= 40 L) or 8^2 points (moments.py:551-554, 589-601);
* image R-family: the same rule on the reflected triangle (image.py:458-464),
  unless the pair is folded (fold band, fold.py:84-118), where the mirror
  element's full-space field takes its place;
* image Q-family: n_q^2 points, n_q = clip(ceil(10 L / sqrt(delta^2 + eps^2))
  + p, 8, 192) with delta the distance to the reflected triangle
  (image.py:211-254); none for a folded pair.

FLOPs per branch are static counts of this repository's kernel source
(csrc/core.hpp, pair.hpp, gen_image.inc), FMA = 2, add/mul/div/sqrt/rcp = 1,
at order P0 in the stress family (the interaction workload):

* far direct, per Gauss point: barycentric point and weight 14, offset and
  R^2 7, 1/sqrt 2, R^-3, R^-5, R^-7 4, in-plane monomials to degree 4 14,
  22 table accumulations 44                                     -> 85
* far image R-family, per Gauss point: 28 + monomials to degree 5 20 +
  R^-3..R^-9 5 + 49 table accumulations 98                      -> 151
* Q-family, contracted (interaction, one or two outputs), per Gauss point:
  barycentric point, weight and offsets 26 and the FP64 accumulation 2 (FP64
  in both modes); R, 1/R, 1/Q 10; the blocked monomial sum of
  csrc/gen_image.inc q_blocked: 44 chain products, 22 first-use / Horner
  products, 166 FMAs -> 398                                      -> 436
  (in the mixed mode 28 of them FP64, 408 FP32); per pair: output functional
  G 160, kappa 1134 FMAs 2268, budget-law order (triangle distance) ~170
                                                                 -> 2600
* per pair, closed-form near direct (three edges, solid angle, recursion)
  ~2.6e3; near image ~5.2e3; direct lift + contraction + rotation ~3.0e3;
  R-family lift + records + contraction ~6.0e3; epilogue 2e2.

The contracted interaction (one or two outputs; k_far + k_near) evaluates
the far field pointwise instead (csrc/kernels.hip):

* per point, FP64 in both modes: barycentric point, weight, offsets 26 and
  the accumulation 2; then R, 1/R 8 (1/Q 2 more), and the series:
  direct (d_blocked: 11 shared + 49 per output)            -> 96 per point
  image R (r_blocked, generated: 71 FMAs, 53 products)     -> 231 per point
  image Q (q_blocked, generated: 166 FMAs, 66 products)    -> 436 per point
  of which FP32 in the mixed mode: 68, 203, 408;
* per pair, FP64: classification, source half F and the block's share of the
  receiver tables ~150; Q kappa 1,002 FMAs 2,004 (not folded); R kappa 432
  FMAs 864 (R far); direct coefficients ~120 (direct far);
* k_near runs the closed forms of the near branches (the constants above).

The near-field constants are averages over the edge regimes; the quadrature
ones are exact per point.  The model is stated at order 0, where a pair is
one column.  At orders 1 and 2 the contracted far pass evaluates each point
once per (row, source) pair for all K node columns (plus K weighted
accumulations per point), so its count is that of n_cols / K pairs; the
histogram's per-column counts overstate it K-fold.
"""
from __future__ import annotations

import numpy as np

from . import defaults, geometry

FLOP_DIRECT_PT = 85
FLOP_IMAGE_PT = 151
FLOP_Q_PT = 436
FLOP_Q_PT_FP64_IN_MIXED = 28
FLOP_Q_PAIR = 2600
FLOP_DIRECT_NEAR = 2600
FLOP_IMAGE_NEAR = 5200
FLOP_DIRECT_LIFT = 3000
FLOP_IMAGE_LIFT = 6000
FLOP_EPILOGUE = 200
# contracted interaction, pointwise far field
FLOP_FAR_D_PT, FLOP_FAR_R_PT, FLOP_FAR_Q_PT = 96, 231, 436
FLOP_FAR_FP64_PT = 28
FLOP_FAR_PAIR, FLOP_FAR_KQ, FLOP_FAR_KR, FLOP_FAR_KD = 150, 2004, 864, 120


# the mixed mode's far pass (csrc/kernels.hip k_far_mixed, mixed_pass): per
# point FP64 15 (offsets 2 x 3 FMAs, weight, the accumulation FMA), FP32 the
# series above (R^2, 1/R and 1/Q included); per (pair, family) kappa formed
# from the receiver tables, 11 FP32 FLOPs per monomial (Q 167, R 72
# monomials); per pair FP64 ~100 (classification, source half) and 120 for
# the direct coefficients of a far pair.
FLOP_MIX_FP64_PT = 15
FLOP_MIX_KQ, FLOP_MIX_KR = 167 * 11, 72 * 11
FLOP_MIX_PAIR = 100


def flops_mixed(h: dict) -> dict:
    """FLOPs of the mixed mode's far pass and near pass (P0, one output) from
    the mixed-rule counts of a histogram (``m_*``).  kappa is counted once
    per (pair, family), the algorithm's minimum; the pass forms it once per
    chunk of points (``m_q_chunks``, ``m_image_chunks``)."""
    pts = h["m_direct_points"] + h["m_image_points"] + h["m_q_points"]
    fp64 = h["m_pairs"] * FLOP_MIX_PAIR + h["m_direct_far"] * FLOP_FAR_KD + pts * FLOP_MIX_FP64_PT
    fp32 = (h["m_direct_points"] * (FLOP_FAR_D_PT - FLOP_FAR_FP64_PT)
            + h["m_image_points"] * (FLOP_FAR_R_PT - FLOP_FAR_FP64_PT)
            + h["m_q_points"] * (FLOP_FAR_Q_PT - FLOP_FAR_FP64_PT)
            + h["m_image_far"] * FLOP_MIX_KR + h["m_q_pairs"] * FLOP_MIX_KQ)
    near = (h["m_direct_near"] * (FLOP_DIRECT_NEAR + FLOP_DIRECT_LIFT + FLOP_EPILOGUE)
            + h["m_image_near"] * (FLOP_IMAGE_NEAR + FLOP_IMAGE_LIFT)
            + h.get("mirror_near", 0) * FLOP_DIRECT_NEAR + h.get("mirror_far12", 0) * 144 * FLOP_DIRECT_PT
            + h.get("mirror_far8", 0) * 64 * FLOP_DIRECT_PT + h["fold_pairs"] * FLOP_DIRECT_LIFT)
    kappa_as_run = (h["m_image_chunks"] * FLOP_MIX_KR + h["m_q_chunks"] * FLOP_MIX_KQ
                    - h["m_image_far"] * FLOP_MIX_KR - h["m_q_pairs"] * FLOP_MIX_KQ)
    return {"far": float(fp64 + fp32), "far_fp32": float(fp32), "far_fp64": float(fp64),
            "near": float(near), "kappa_recomputed": float(kappa_as_run)}


def _tri_dist(obs, v):
    """Vectorised point-triangle distance, obs (N, 3), v (S, 3, 3) -> (N, S)."""
    a, b, c = v[:, 0], v[:, 1], v[:, 2]
    ab, ac = b - a, c - a
    ap = obs[:, None, :] - a[None]
    d20 = np.einsum("nsi,si->ns", ap, ab)
    d21 = np.einsum("nsi,si->ns", ap, ac)
    d00 = np.einsum("si,si->s", ab, ab)
    d01 = np.einsum("si,si->s", ab, ac)
    d11 = np.einsum("si,si->s", ac, ac)
    den = d00 * d11 - d01 * d01
    vv = (d11 * d20 - d01 * d21) / den
    ww = (d00 * d21 - d01 * d20) / den
    uu = 1.0 - vv - ww
    nrm = np.cross(ab, ac)
    plane = np.abs(np.einsum("nsi,si->ns", ap, nrm)) / np.linalg.norm(nrm, axis=1)
    best = np.full(plane.shape, np.inf)
    for p0, p1 in ((a, b), (b, c), (c, a)):
        e = p1 - p0
        t = np.clip(np.einsum("nsi,si->ns", obs[:, None, :] - p0[None], e) / np.einsum("si,si->s", e, e), 0, 1)
        q = p0[None] + t[..., None] * e[None]
        best = np.minimum(best, np.linalg.norm(obs[:, None, :] - q, axis=2))
    inside = (uu >= 0) & (vv >= 0) & (ww >= 0)
    return np.where(inside, plane, best)


def histogram(obs, tris, eps, order=0, fold=True, chunk=256) -> dict:
    """Branch counts over all (observer, source) pairs."""
    obs = np.asarray(obs, float)
    recs = geometry.source_records(tris, np.broadcast_to(eps, (tris.shape[0],)))
    c = recs[:, geometry.S_C:geometry.S_C + 3]
    ic = recs[:, geometry.S_IC:geometry.S_IC + 3]
    L = recs[:, geometry.S_L]
    e = recs[:, geometry.S_EPS]
    vi = np.asarray(tris, float) * np.array([1.0, 1.0, -1.0])
    strad = recs[:, geometry.S_STRADDLE] != 0.0
    crease = recs[:, geometry.S_CREASE:geometry.S_CREASE + 2]
    h = dict(pairs=0, direct_near=0, direct_far12=0, direct_far8=0, image_near=0,
             image_far12=0, image_far8=0, fold_pairs=0, q_points=0, q_pairs=0)
    for lo in range(0, obs.shape[0], chunk):
        o = obs[lo:lo + chunk]
        band = np.zeros((o.shape[0], tris.shape[0]), bool)
        if fold:
            reach = np.hypot(o[:, None, 0] - crease[None, :, 0], o[:, None, 1] - crease[None, :, 1])
            band = (strad[None] & (-o[:, None, 2] < defaults.FOLD_BAND_DEPTH_EPS * e[None])
                    & (reach < defaults.FOLD_REACH_EPS * e[None]))
        for cen, key in ((c, "direct"), (ic, "image")):
            D = np.sqrt(np.sum((o[:, None, :] - cen[None]) ** 2, axis=2) + e[None] ** 2)
            near = D <= defaults.D_STAR * L[None]
            far8 = D > defaults.D_STAR_DISTANT * L[None]
            sel = ~band if key == "image" else np.ones_like(band)
            h[f"{key}_near"] += int(np.sum(near & sel))
            h[f"{key}_far8"] += int(np.sum(~near & far8 & sel))
            h[f"{key}_far12"] += int(np.sum(~near & ~far8 & sel))
        delta = _tri_dist(o, vi)
        nq = np.clip(np.ceil(defaults.IMAGE_Q_BUDGET_C * L[None] / np.sqrt(delta ** 2 + e[None] ** 2)) + order,
                     defaults.IMAGE_Q_GAUSS_MIN, defaults.IMAGE_Q_GAUSS_MAX)
        h["q_points"] += int(np.sum(np.where(band, 0, nq * nq)))
        h["q_pairs"] += int(np.sum(~band))
        h["fold_pairs"] += int(np.sum(band))
        h["pairs"] += band.size
    return h


def flops(h: dict) -> dict:
    """FLOPs per pass from a histogram (P0, stress family)."""
    pair = (h["direct_near"] * FLOP_DIRECT_NEAR + h["direct_far12"] * 144 * FLOP_DIRECT_PT
            + h["direct_far8"] * 64 * FLOP_DIRECT_PT + h["pairs"] * (FLOP_DIRECT_LIFT + FLOP_EPILOGUE)
            + h["image_near"] * FLOP_IMAGE_NEAR + h["image_far12"] * 144 * FLOP_IMAGE_PT
            + h["image_far8"] * 64 * FLOP_IMAGE_PT
            + h.get("mirror_near", 0) * FLOP_DIRECT_NEAR + h.get("mirror_far12", 0) * 144 * FLOP_DIRECT_PT
            + h.get("mirror_far8", 0) * 64 * FLOP_DIRECT_PT + h["fold_pairs"] * FLOP_DIRECT_LIFT
            + (h["pairs"] - h["fold_pairs"]) * FLOP_IMAGE_LIFT)
    q = h["q_points"] * FLOP_Q_PT + h["q_pairs"] * FLOP_Q_PAIR
    q_fp32_mixed = h["q_points"] * (FLOP_Q_PT - FLOP_Q_PT_FP64_IN_MIXED)
    # contracted interaction: far pass and near pass
    d_pts = h["direct_far12"] * 144 + h["direct_far8"] * 64
    r_pts = h["image_far12"] * 144 + h["image_far8"] * 64
    d_far = h["direct_far12"] + h["direct_far8"]
    r_far = h["image_far12"] + h["image_far8"]
    far = (d_pts * FLOP_FAR_D_PT + r_pts * FLOP_FAR_R_PT + h["q_points"] * FLOP_FAR_Q_PT
           + h["pairs"] * FLOP_FAR_PAIR + h["q_pairs"] * FLOP_FAR_KQ + r_far * FLOP_FAR_KR
           + d_far * FLOP_FAR_KD)
    far_fp32 = (d_pts * (FLOP_FAR_D_PT - FLOP_FAR_FP64_PT) + r_pts * (FLOP_FAR_R_PT - FLOP_FAR_FP64_PT)
                + h["q_points"] * (FLOP_FAR_Q_PT - FLOP_FAR_FP64_PT))
    near = (h["direct_near"] * (FLOP_DIRECT_NEAR + FLOP_DIRECT_LIFT + FLOP_EPILOGUE)
            + h["image_near"] * (FLOP_IMAGE_NEAR + FLOP_IMAGE_LIFT)
            + h.get("mirror_near", 0) * FLOP_DIRECT_NEAR + h.get("mirror_far12", 0) * 144 * FLOP_DIRECT_PT
            + h.get("mirror_far8", 0) * 64 * FLOP_DIRECT_PT + h["fold_pairs"] * FLOP_DIRECT_LIFT)
    return {"pair": float(pair), "q": float(q), "q_fp32_in_mixed": float(q_fp32_mixed),
            "far": float(far), "far_fp32_in_mixed": float(far_fp32), "near": float(near)}


# The fold's point kernels (k_fold: the Boussinesq surface integral of the
# folded pairs; per folded (pair, node column) every polar-grid point),
# static counts of csrc/fold.hpp and core.hpp at order 0, FMA = 2:
# * per point, both kinds: the grid point (radius, angle, cos, sin, weight)
#   ~20, the Boussinesq point stress ~40, the load minus its centre value and
#   the 18 weighted accumulations ~45                               -> 105
# * the contraction of one slip direction's surface load (direct_contract:
#   the E.F products 72, three local quadratics 120, the series over the W
#   tables 220, the rest ~13)                                      -> 425
# * near point (closed form): the weighted tables (FLOP_DIRECT_NEAR) + three
#   contractions + the point                                         -> 3,980
# * far point, FP64 rule: the frame offsets ~30 + three contractions + the
#   point, plus per Gauss point the table accumulation FLOP_DIRECT_PT
# * far point, mixed rule: the point (FP64) plus per Gauss point the direct
#   series for three slip outputs (FP32: R, 1/R 8, shared 11, 49 per output;
#   FP64: the point, weight and offsets 26, three accumulations 6 -> 164 FP32 + 32 FP64)
FLOP_FOLD_POINT = 105
FLOP_FOLD_CONTRACT = 425
FLOP_FOLD_NEAR_PT = FLOP_DIRECT_NEAR + 3 * FLOP_FOLD_CONTRACT + FLOP_FOLD_POINT
FLOP_FOLD_FAR_PT = 30 + 3 * FLOP_FOLD_CONTRACT + FLOP_FOLD_POINT
FLOP_FOLD_MIX_GP32, FLOP_FOLD_MIX_GP64 = 164, 32


def flops_fold(fh: dict, mixed: bool) -> dict:
    """FLOPs of the fold's point kernels from Prepared.fold_histogram
    counts (estimated per point from the static counts above)."""
    near = fh["fold_near_points"] * FLOP_FOLD_NEAR_PT
    if mixed:
        fp32 = fh["fold_far_gauss_points"] * FLOP_FOLD_MIX_GP32
        far64 = fh["fold_far_points"] * FLOP_FOLD_POINT + fh["fold_far_gauss_points"] * FLOP_FOLD_MIX_GP64
    else:
        fp32 = 0.0
        far64 = fh["fold_far_points"] * FLOP_FOLD_FAR_PT + fh["fold_far_gauss_points"] * FLOP_DIRECT_PT
    return {"near": float(near), "far_fp64": float(far64), "far_fp32": float(fp32),
            "fp64": float(near + far64), "fp32": float(fp32)}
