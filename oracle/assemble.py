"""Per-source assembly of every matrix form (oracle).

Restates src/mhs/kernels/assemble.py:69-366 and the entry points of
src/mhs/matrices.py:231-580 on already-validated inputs: ``obs`` (n_obs, 3),
``tris`` (n_src, 3, 3), ``eps_arr`` (n_src,), ``ids`` None (discontinuous) or
(n_src, K) shared column ids (continuous, scatter-add).
"""
from __future__ import annotations

import numpy as np

from . import fold as _fold
from .fullspace import blocks
from .geometry import n_nodes, unit_normal
from .image import image_total


def _put(out, s, K, block, ids, axis):
    sl = [slice(None)] * out.ndim
    if ids is None:
        sl[axis] = slice(s * K, (s + 1) * K)
        out[tuple(sl)] = block
    else:
        sl[axis] = ids[s]
        out[tuple(sl)] += block


def _add(out, s, K, block, ids, axis):
    sl = [slice(None)] * out.ndim
    sl[axis] = slice(s * K, (s + 1) * K) if ids is None else ids[s]
    out[tuple(sl)] += block


def image_half(obs, tri, mu, lam, eps, order, want, fold):
    """Lift everywhere, folded image in the band (assemble.py:108-120)."""
    img = image_total(obs, tri, order, float(mu), float(lam), float(eps), want=tuple(want))
    if fold and _fold.straddles(tri, eps):
        sel = _fold.band(obs, tri, eps)
        if sel.size:
            fb = _fold.image(obs[sel], tri, mu, lam, eps, order, tuple(want))
            for k in want:
                img[k][sel] = fb[k]
    return img


def hooke_normal(n, mu, lam):
    """C:sym(slip (x) n) per unit slip k, (3, 3, 3) (assemble.py:123-128)."""
    eye = np.eye(3)
    return (lam * np.einsum("mn,k->mnk", eye, n)
            + mu * (np.einsum("mk,n->mnk", eye, n) + np.einsum("nk,m->mnk", eye, n)))


def eigen_terms(obs, tri, mu, lam, eps, order, phi, fold):
    """[(phi (n_obs, K), C (3, 3, 3))] (assemble.py:131-149)."""
    terms = [(np.asarray(phi), hooke_normal(unit_normal(tri), mu, lam))]
    if fold and _fold.straddles(tri, eps):
        sel = _fold.band(obs, tri, eps)
        if sel.size:
            trim = np.asarray(tri, float) * _fold.REFLECT
            phi_m = np.zeros_like(terms[0][0])
            phi_m[sel] = blocks(obs[sel], trim, mu, lam, eps, ("E",), order)["E"]
            terms.append((phi_m, hooke_normal(unit_normal(trim), mu, lam)
                           * _fold.MIRROR_SLIP_SIGN[None, None, :]))
    return terms


def disp(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True):
    """assemble.py:152-208: direct pass then image pass."""
    K = n_nodes(order)
    for s in range(tris.shape[0]):
        U = blocks(obs, tris[s], mu, lam, eps_arr[s], ("U",), order)["U"]
        _put(out, s, K, np.moveaxis(U, 1, 2), ids, 2)
    for s in range(tris.shape[0]):
        img = image_half(obs, tris[s], mu, lam, eps_arr[s], order, ("U",), fold)
        _add(out, s, K, np.moveaxis(img["U"], 1, 2), ids, 2)
    return out


def total_stress(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True):
    """assemble.py:173-230."""
    K = n_nodes(order)
    for s in range(tris.shape[0]):
        H = blocks(obs, tris[s], mu, lam, eps_arr[s], ("H",), order)["H"]
        _put(out, s, K, np.moveaxis(H, 1, 3), ids, 3)
    for s in range(tris.shape[0]):
        img = image_half(obs, tris[s], mu, lam, eps_arr[s], order, ("H",), fold)
        _add(out, s, K, np.moveaxis(img["H"], 1, 3), ids, 3)
    return out


def eigenstress(obs, tris, mu, lam, eps_arr, out, order=0, ids=None, fold=True):
    """assemble.py:334-366."""
    K = n_nodes(order)
    for s in range(tris.shape[0]):
        phi = blocks(obs, tris[s], mu, lam, eps_arr[s], ("E",), order)["E"]
        block = None
        for ph, C in eigen_terms(obs, tris[s], mu, lam, eps_arr[s], order, phi, fold):
            t = ph[:, None, None, :, None] * C[None, :, :, None, :]
            block = t if block is None else block + t
        _put(out, s, K, block, ids, 3)
    return out


def _stress_block(obs, tri, mu, lam, eps, want, order, fold):
    """Direct + image total stress and eigen terms (assemble.py:249-262)."""
    blk = blocks(obs, tri, mu, lam, eps, want, order)
    img = image_half(obs, tri, mu, lam, eps, order, ("H",), fold)
    terms = eigen_terms(obs, tri, mu, lam, eps, order, blk["E"], fold) if "E" in want else None
    return blk["H"] + img["H"], terms


def _eig_traction(terms, nrm):
    """assemble.py:233-246."""
    out = None
    for phi, C in terms:
        t = phi[:, :, None, None] * np.einsum("ijk,oj->oik", C, nrm)[:, None, :, :]
        out = t if out is None else out + t
    return out


def traction(obs, nrm, tris, mu, lam, eps_arr, out, subtract=True, order=0, ids=None, fold=True):
    """assemble.py:265-292."""
    want = ("H", "E") if subtract else ("H",)
    K = n_nodes(order)
    for s in range(tris.shape[0]):
        H, terms = _stress_block(obs, tris[s], mu, lam, eps_arr[s], want, order, fold)
        t = np.einsum("okijm,oj->okim", H, nrm)
        if subtract:
            t = t - _eig_traction(terms, nrm)
        _put(out, s, K, np.moveaxis(t, 1, 2), ids, 2)
    return out


def interaction(obs, nrm, rec_basis, tris, src_basis, mu, lam, eps_arr, out,
                subtract=True, order=0, ids=None, fold=True):
    """assemble.py:295-331."""
    want = ("H", "E") if subtract else ("H",)
    K = n_nodes(order)
    for s in range(tris.shape[0]):
        H, terms = _stress_block(obs, tris[s], mu, lam, eps_arr[s], want, order, fold)
        t = np.einsum("okijm,oj->okim", H, nrm)
        if subtract:
            t = t - _eig_traction(terms, nrm)
        if ids is None:
            out[:, s * K:(s + 1) * K] = np.einsum("oai,okim,bm->okab", rec_basis, t, src_basis[s])
        else:
            out[:, ids[s]] += np.einsum("oai,okim,kbm->okab", rec_basis, t, src_basis[ids[s]])
    return out
