"""The reference's assembly seam (src/mhs/kernels/assemble.py:186-366) served
by the device kernels: each ``assemble_*`` called the way the reference's
public functions call it -- contiguous float64 arrays, plain floats, a
zeroed caller-allocated ``out``, ``ids`` for shared columns -- against the
reference's golden matrices at the all-FP64 bound, 1e-12 relmax."""
import numpy as np
import pytest

from conftest import relmax as _relmax

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relmax(a, b):
    r = _relmax(a, b)
    print(f"  relmax {r:.3e}", flush=True)
    return r


@pytest.fixture(scope="module")
def S():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mhs_amd import seam
    return seam


def _frames(tris, K):
    from mhs_amd import tdcs
    return np.repeat(tdcs.slip_frame(tris), K, axis=0)


def test_seam_discontinuous_forms(S, small):
    import mhs_amd as M
    g = small
    mu, lam, tb, eb, ob = float(g["mu"]), float(g["lam"]), g["tb"], g["eb"], g["ob"]
    nrm = g["nrm_b"] / np.linalg.norm(g["nrm_b"], axis=1, keepdims=True)
    n, m = tb.shape[0], ob.shape[0]
    for p in (0, 1, 2):
        K = (p + 1) * (p + 2) // 2
        print(f"case seam disp / traction / interaction p{p}", flush=True)
        out = np.zeros((m, 3, n * K, 3))
        assert S.assemble_halfspace_disp(ob, tb, mu, lam, eb, out, order=p) is out
        assert relmax(out, g[f"disp_b_p{p}"]) < TOL
        out = np.zeros((m, 3, n * K, 3))
        S.assemble_halfspace_traction(ob, nrm, tb, mu, lam, eb, out, order=p)
        assert relmax(out, g[f"trac_b_p{p}"]) < TOL
        pts = M.collocation_points(tb, p)
        fr = _frames(tb, K)
        src = M.tdcs.slip_frame(tb)
        out = np.zeros((n * K, n * K, 3, 3))
        S.assemble_interaction(pts, fr[:, 2, :], fr, tb, src, mu, lam, eb, out, order=p)
        assert relmax(out, g[f"inter_b_p{p}"]) < TOL
    print("case seam total stress / eigenstress / traction without eigenstress", flush=True)
    out = np.zeros((m, 3, 3, n, 3))
    S.assemble_halfspace_total_stress(ob, tb, mu, lam, eb, out)
    assert relmax(out, g["total_b_p0"]) < TOL
    eig = np.zeros((m, 3, 3, n, 3))
    S.assemble_eigenstress(ob, tb, mu, lam, eb, eig)
    assert relmax(eig, g["eig_b_p0"]) < TOL
    # the reference forms the elastic stress as total - eigenstress
    assert relmax(out - eig, g["stress_b_p0"]) < TOL
    out = np.zeros((m, 3, n, 3))
    S.assemble_halfspace_traction(ob, nrm, tb, mu, lam, eb, out, subtract_eigenstress=False)
    assert relmax(out, g["trac_b_noeig"]) < TOL
    print("case seam float32 out", flush=True)
    o32 = np.zeros((m, 3, n, 3), np.float32)
    S.assemble_halfspace_disp(ob, tb, mu, lam, eb, o32)
    assert relmax(o32, g["disp_b_p0"]) < 1e-6


def test_seam_surface_fold(S, small):
    g = small
    mu, lam, ts, os_ = float(g["mu"]), float(g["lam"]), g["ts"], g["os"]
    es = np.full(ts.shape[0], float(g["es"]))
    n, m = ts.shape[0], os_.shape[0]
    print("case seam surface disp with and without the fold", flush=True)
    for fold, key in ((True, "disp_s_p0"), (False, "disp_s_nofold")):
        out = np.zeros((m, 3, n, 3))
        S.assemble_halfspace_disp(os_, ts, mu, lam, es, out, fold=fold)
        assert relmax(out, g[key]) < TOL
    eig = np.zeros((m, 3, 3, n, 3))
    S.assemble_eigenstress(os_, ts, mu, lam, es, eig)
    assert relmax(eig, g["eig_s_p0"]) < TOL


def test_seam_shared_columns_add(S, small):
    """ids: element nodes scatter-ADD into shared columns of a zeroed out
    (the continuous layout's assembly side, assemble.py:58-76): traction
    against the golden continuous matrix, the interaction with a per-column
    source basis, and a second call adding the same matrix again."""
    import mhs_amd as M
    g = small
    mu, lam, tb, eb, ob = float(g["mu"]), float(g["lam"]), g["tb"], g["eb"], g["ob"]
    nrm = g["nrm_b"] / np.linalg.norm(g["nrm_b"], axis=1, keepdims=True)
    sn = M.shared_nodes(tb)
    ids = np.asarray(sn.ids)
    print("case seam traction, shared columns", flush=True)
    out = np.zeros((ob.shape[0], 3, sn.n, 3))
    S.assemble_halfspace_traction(ob, nrm, tb, mu, lam, eb, out, order=2, ids=ids)
    assert relmax(out, g["trac_b_cont"]) < TOL
    once = out.copy()
    S.assemble_halfspace_traction(ob, nrm, tb, mu, lam, eb, out, order=2, ids=ids)
    assert np.array_equal(out, 2.0 * once)
    print("case seam interaction, shared columns, per-column source basis", flush=True)
    fr = sn.frames
    out = np.zeros((sn.n, sn.n, 3, 3))
    S.assemble_interaction(sn.coords, fr[:, 2, :], fr, tb, fr, mu, lam, eb, out, order=2, ids=ids)
    assert relmax(out, g["inter_b_cont"]) < TOL
    print("case seam disp, shared columns", flush=True)
    out = np.zeros((ob.shape[0], 3, sn.n, 3))
    S.assemble_halfspace_disp(ob, tb, mu, lam, eb, out, order=2, ids=ids)
    assert relmax(out, g["disp_b_cont"]) < TOL
