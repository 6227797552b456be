"""GPU parity: the HIP kernels (through the C-ABI) against the reference's
golden vectors and the FP64 oracle, plus the reference's own identities
rerun on GPU output (tests/gates/mhs/verify_image_kernel.py:439-645,
verify_fold.py:242-262, verify_shared_nodes.py:332-442).

Tolerance: relmax (max|a - b| / max|b|) <= 1e-12 in FP64, the all-FP64
parity bound; every comparison prints its relmax.  Identities the reference
states bitwise (total - eigenstress == stress, parts=True == the separate
calls) are asserted bitwise; uniform nodal slip at P1/P2 reproducing P0 at
1e-11 at every entry point, as verify_image_kernel.py:439-534 states it.
"""
import numpy as np
import pytest

from conftest import GOLDEN
from conftest import relmax as _relmax

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relmax(a, b):
    """relmax, printed (the case name is printed before each group)."""
    r = _relmax(a, b)
    print(f"  relmax {r:.3e}", flush=True)
    return r


@pytest.fixture(scope="module")
def M():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mhs_amd
    return mhs_amd


def _mat(M, g):
    return M.Material(float(g["mu"]), float(g["lam"]))


@pytest.mark.parametrize("p", [0, 1, 2])
def test_buried_forms(M, small, p):
    g = small
    mat, tb, eb, ob = _mat(M, g), g["tb"], g["eb"], g["ob"]
    print(f"case buried p{p}", flush=True)
    assert relmax(M.disp_matrix(ob, tb, mat, eb, order=p), g[f"disp_b_p{p}"]) < TOL
    sig, eig = M.stress_matrix(ob, tb, mat, eb, order=p, parts=True)
    assert relmax(sig, g[f"stress_b_p{p}"]) < TOL
    assert relmax(eig, g[f"eig_b_p{p}"]) < TOL
    assert relmax(M.traction_matrix(ob, g["nrm_b"], tb, mat, eb, order=p), g[f"trac_b_p{p}"]) < TOL
    assert relmax(M.interaction_matrix(tb, mat, eb, order=p), g[f"inter_b_p{p}"]) < TOL


def test_buried_variants(M, small):
    g = small
    mat, tb, eb, ob = _mat(M, g), g["tb"], g["eb"], g["ob"]
    print("case buried variants", flush=True)
    assert relmax(M.total_stress_matrix(ob, tb, mat, eb), g["total_b_p0"]) < TOL
    assert relmax(M.elastic_strain_matrix(ob, tb, mat, eb), g["strain_b_p0"]) < TOL
    assert relmax(M.traction_matrix(ob, g["nrm_b"], tb, mat, eb, subtract_eigenstress=False),
                  g["trac_b_noeig"]) < TOL
    assert relmax(M.interaction_matrix(tb, mat, eb, receiver="strike", source="strike"), g["inter_b_ss"]) < TOL
    assert relmax(M.interaction_matrix(tb, mat, eb, order=1, rows=(5, 17), receiver=("dip", "normal"),
                                       source="tensile"), g["inter_b_rows"]) < TOL
    assert relmax(M.interaction_matrix(tb, mat, eb, obs_tris=tb[:3], shrink=0.2, order=1),
                  g["inter_b_obs"]) < TOL


def test_continuous_layout(M, small):
    g = small
    mat, tb, eb, ob = _mat(M, g), g["tb"], g["eb"], g["ob"]
    print("case continuous", flush=True)
    assert relmax(M.interaction_matrix(tb, mat, eb, order=2, layout="continuous"), g["inter_b_cont"]) < TOL
    assert relmax(M.disp_matrix(ob, tb, mat, eb, order=2, layout="continuous"), g["disp_b_cont"]) < TOL
    assert relmax(M.traction_matrix(ob, g["nrm_b"], tb, mat, eb, order=2, layout="continuous"),
                  g["trac_b_cont"]) < TOL


def test_surface_fold(M, small):
    """Surface-reaching sources: folded mirror + Boussinesq in the band."""
    g = small
    mat, ts, es, os_ = _mat(M, g), g["ts"], float(g["es"]), g["os"]
    print("case surface fold", flush=True)
    for p in (0, 2):
        assert relmax(M.disp_matrix(os_, ts, mat, es, order=p), g[f"disp_s_p{p}"]) < TOL
        assert relmax(M.traction_matrix(os_, g["nrm_s"], ts, mat, es, order=p), g[f"trac_s_p{p}"]) < TOL
        assert relmax(M.interaction_matrix(ts, mat, es, order=p), g[f"inter_s_p{p}"]) < TOL
    sig, eig = M.stress_matrix(os_, ts, mat, es, parts=True)
    assert relmax(sig, g["stress_s_p0"]) < TOL
    assert relmax(eig, g["eig_s_p0"]) < TOL
    assert relmax(M.disp_matrix(os_, ts, mat, es, fold=False), g["disp_s_nofold"]) < TOL
    assert relmax(M.stress_matrix(os_, ts, mat, es, fold=False), g["stress_s_nofold"]) < TOL
    assert relmax(M.interaction_matrix(ts, mat, es, order=2, layout="continuous", receiver="strike",
                                       source=("strike", "dip")), g["inter_s_cont"]) < TOL


def test_c1_rows(M):
    from mhs_amd import workloads
    obs, tris, eps = workloads.c1_disp()
    g = np.load(GOLDEN / "c1_rows.npz")
    print("case C1 rows", flush=True)
    G = M.disp_matrix(obs[g["sel"]], tris, M.Material(30.0, 30.0), eps)
    assert relmax(G, g["G"]) < TOL


def test_c3_rows(M):
    """Row blocks of the n = 10k strike/strike interaction (fold active)."""
    from mhs_amd import workloads
    tris, eps = workloads.c3_interaction()
    g = np.load(GOLDEN / "c3_rows.npz")
    print("case C3 rows", flush=True)
    for lo in (0, 5000, 5024, 9998):
        K = M.interaction_matrix(tris, M.Material(30.0, 30.0), eps, receiver="strike", source="strike",
                                 rows=(lo, lo + 2))
        assert relmax(K, g[f"rows_{lo}"]) < TOL, lo


def test_random_vs_oracle(M):
    """Seeded random meshes and observers against the FP64 oracle."""
    from mhs_amd import workloads
    from oracle import api as O
    rng = np.random.default_rng(1234)
    for case in range(3):
        print(f"case random {case}", flush=True)
        P = workloads.plane(rng.uniform(0.8, 2.0), rng.uniform(0.6, 1.5), 2, 1, rng.uniform(20, 90),
                            -rng.uniform(0.0, 0.4), 0.0, 0.0, rng.uniform(0, 180))
        tris = workloads.orient_up(workloads.grid_mesh(P + rng.uniform(-0.03, 0.03, P.shape) * [1, 1, 0]))
        tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
        eps = rng.uniform(0.05, 0.15, tris.shape[0])
        obs = np.column_stack([rng.uniform(-3, 3, 12), rng.uniform(-3, 3, 12), -rng.uniform(0, 2, 12)])
        mat = M.Material(30.0, float(rng.uniform(10, 60)))
        p = case
        assert relmax(M.disp_matrix(obs, tris, mat, eps, order=p),
                      O.disp_matrix(obs, tris, mat.mu, mat.lam, eps, order=p)) < TOL
        assert relmax(M.interaction_matrix(tris, mat, eps, order=p, receiver="dip", source=("strike", "dip")),
                      O.interaction_matrix(tris, mat.mu, mat.lam, eps, order=p, receiver="dip",
                                           source=("strike", "dip"))) < TOL


def test_node_columns_share_points(M):
    """The contracted far and near passes evaluate a (receiver, source) pair
    once for all K node columns: one and two outputs at orders 1 and 2
    against the FP64 oracle on a buried 4-triangle fault (near and far pairs,
    no fold: the oracle's fold is slow), and the mixed mode (FP32 points and
    node weights) against FP64 on a 240-triangle fault reaching the surface."""
    from mhs_amd import workloads
    from oracle import api as O
    rng = np.random.default_rng(4321)
    P = workloads.plane(1.6, 1.0, 2, 1, 70.0, -0.6, 0.0, 0.0, 30.0)
    tris = workloads.orient_up(workloads.grid_mesh(P + rng.uniform(-0.03, 0.03, P.shape) * [1, 1, 0]))
    tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
    eps = rng.uniform(0.05, 0.15, tris.shape[0])
    mat = M.Material(30.0, 20.0)
    for p in (1, 2):
        for rec, srcb in (("strike", "strike"), ("normal", ("strike", "dip"))):
            print(f"case nodes p{p} {rec} {srcb}", flush=True)
            assert relmax(M.interaction_matrix(tris, mat, eps, order=p, receiver=rec, source=srcb),
                          O.interaction_matrix(tris, mat.mu, mat.lam, eps, order=p, receiver=rec,
                                               source=srcb)) < TOL
    Pb = workloads.plane(30.0, 8.0, 12, 10, 60.0, 0.0)
    tb = workloads.orient_up(workloads.grid_mesh(Pb))
    for p in (1, 2):
        print(f"case nodes mixed p{p}", flush=True)
        K64 = M.interaction_matrix(tb, mat, 0.1, order=p, receiver="strike", source="strike")
        with M.precision("mixed"):
            K32 = M.interaction_matrix(tb, mat, 0.1, order=p, receiver="strike", source="strike")
        err = relmax(K32, K64)
        print(f"mixed relmax vs fp64 p{p} {err:.3e}", flush=True)
        assert err < 1e-6


def test_block_drivers_equal_one_call(M):
    """The reference-compatible block drivers (parallel.by_source /
    by_rows) stitch device blocks into bitwise the single call's matrix:
    traction by source blocks, the contracted interaction by source blocks
    with the receivers bound, and the continuous layout by row blocks."""
    import functools
    from mhs_amd import parallel as mp, workloads
    rng = np.random.default_rng(99)
    Pl = workloads.plane(12.0, 6.0, 8, 4, 55.0, 0.0)
    tris = workloads.orient_up(workloads.grid_mesh(Pl + rng.uniform(-0.05, 0.05, Pl.shape) * [1, 1, 0]))
    tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
    mat = M.Material(30.0, 30.0)
    obs = np.column_stack([rng.uniform(-8, 8, 50), rng.uniform(-8, 8, 50), -rng.uniform(0, 6, 50)])
    nrm = rng.normal(size=(50, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    print("case by_source traction", flush=True)
    call = functools.partial(M.traction_matrix, obs, nrm, material=mat, eps=0.1, order=1)
    assert np.array_equal(mp.by_source(call, tris, block_bytes=200_000), call(tris))
    print("case by_source interaction", flush=True)
    call = functools.partial(M.interaction_matrix, material=mat, eps=0.1, obs_tris=tris,
                             receiver="strike", source="strike")
    assert np.array_equal(mp.by_source(call, tris, source_axis=1, tris_kw="tris", block_bytes=20_000),
                          call(tris=tris))
    print("case by_rows continuous", flush=True)
    sn = M.shared_nodes(tris)
    call = functools.partial(M.interaction_matrix, tris, mat, 0.1, order=2, layout="continuous", shared=sn)
    full = call()
    assert np.array_equal(mp.by_rows(call, full.shape[0], block_bytes=100_000), full)


def test_row_chunks_beyond_one_launch(M):
    """More receiver rows than one launch takes (32,768): the launch loop's
    row offsets, for the pair, Q-family, far, near and fold kernels.  Rows on
    both sides of the chunk boundary against the FP64 oracle; the whole
    matrix bitwise against row-streamed blocks (other launch geometry)."""
    from mhs_amd import parallel as mp, workloads
    from oracle import api as O
    rng = np.random.default_rng(21)
    Pl = workloads.plane(2.0, 1.0, 2, 1, 60.0, 0.0)
    tris = workloads.orient_up(workloads.grid_mesh(Pl + rng.uniform(-0.02, 0.02, Pl.shape) * [1, 1, 0]))
    tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
    mat = M.Material(30.0, 30.0)
    n = 40_000
    obs = np.column_stack([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n), -rng.uniform(0, 1.5, n)])
    obs[32_760:32_780, 2] = -rng.uniform(0, 0.3, 20)            # fold band around the boundary
    pick = np.r_[0, 32_766:32_770, n - 1]
    print("case row chunks traction", flush=True)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    T = M.traction_matrix(obs, nrm, tris, mat, 0.1, order=1)
    assert relmax(T[pick], O.traction_matrix(obs[pick], nrm[pick], tris, mat.mu, mat.lam, 0.1, order=1)) < TOL
    assert np.array_equal(mp.stream_rows(M.traction_matrix, obs, nrm, tris, mat, 0.1, order=1,
                                         block_bytes=T[0].nbytes * 9000), T)
    print("case row chunks contracted interaction", flush=True)
    otris = np.repeat(tris[:1], n, axis=0)
    kw = dict(receiver="strike", source="strike", obs_tris=otris, obs_pts=obs)
    K = M.interaction_matrix(tris, mat, 0.1, **kw)
    ref = O.interaction_matrix(tris, mat.mu, mat.lam, 0.1, receiver="strike", source="strike",
                               obs_tris=otris[pick], obs_pts=obs[pick])
    assert relmax(K[pick], ref) < TOL
    with M.precision("mixed"):
        Km = M.interaction_matrix(tris, mat, 0.1, **kw)
    assert relmax(Km, K) < 1e-6


def test_stream_rows_equals_one_call(M, tmp_path):
    """Row blocks with overlapped device-to-host copies land bitwise the
    direct call's matrix: displacement, stress, traction, the contracted and
    the continuous interaction; a memmap and a float32 out=."""
    from mhs_amd import parallel as mp, workloads
    rng = np.random.default_rng(5)
    Pl = workloads.plane(10.0, 5.0, 8, 4, 50.0, 0.0)
    tris = workloads.orient_up(workloads.grid_mesh(Pl + rng.uniform(-0.05, 0.05, Pl.shape) * [1, 1, 0]))
    tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
    mat = M.Material(30.0, 30.0)
    obs = np.column_stack([rng.uniform(-8, 8, 70), rng.uniform(-8, 8, 70), -rng.uniform(0, 5, 70)])
    nrm = rng.normal(size=(70, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cases = [(M.disp_matrix, (obs, tris, mat, 0.1), {"order": 1}),
             (M.stress_matrix, (obs, tris, mat, 0.1), {}),
             (M.traction_matrix, (obs, nrm, tris, mat, 0.1), {"order": 2}),
             (M.interaction_matrix, (tris, mat, 0.1), {"receiver": "strike", "source": "strike"}),
             (M.interaction_matrix, (tris, mat, 0.1), {"order": 2, "layout": "continuous",
                                                       "shared": M.shared_nodes(tris)})]
    for fn, a, kw in cases:
        print(f"case stream_rows {fn.__name__} {kw.get('order', 0)}", flush=True)
        want = fn(*a, **kw)
        assert np.array_equal(mp.stream_rows(fn, *a, block_bytes=want[0].nbytes * 7, **kw), want)
    want = M.traction_matrix(obs, nrm, tris, mat, 0.1, order=2)
    mm = np.lib.format.open_memmap(tmp_path / "t.npy", mode="w+", dtype=np.float64, shape=want.shape)
    mp.stream_rows(M.traction_matrix, obs, nrm, tris, mat, 0.1, order=2, out=mm, block_bytes=want[0].nbytes * 9)
    mm.flush()
    assert np.array_equal(np.load(tmp_path / "t.npy"), want)
    o32 = np.zeros(want.shape, np.float32)
    mp.stream_rows(M.traction_matrix, obs, nrm, tris, mat, 0.1, order=2, out=o32)
    assert np.array_equal(o32, want.astype(np.float32))
    with pytest.raises(ValueError, match="parts=True"):
        mp.stream_rows(M.stress_matrix, obs, tris, mat, 0.1, parts=True)


def test_public_out_streams_beyond_one_block(M, tmp_path, monkeypatch):
    """An out= larger than one device block (matrices.py:169-189: out= may
    be a memmap larger than memory) is filled by receiver-row blocks, not
    assembled whole on the device: bitwise the one-call matrix, into a
    memmap and a float32 array."""
    from mhs_amd import defaults, parallel as mp, workloads
    rng = np.random.default_rng(8)
    tris = workloads.orient_up(workloads.grid_mesh(workloads.plane(10.0, 5.0, 6, 3, 60.0, -0.5)))
    mat = M.Material(30.0, 30.0)
    obs = np.column_stack([rng.uniform(-8, 8, 50), rng.uniform(-8, 8, 50), -rng.uniform(0, 5, 50)])
    nrm = rng.normal(size=(50, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    want = M.traction_matrix(obs, nrm, tris, mat, 0.1, order=1)
    wint = M.interaction_matrix(tris, mat, 0.1, order=1)
    calls = []
    real = mp.stream_job
    monkeypatch.setattr(mp, "stream_job", lambda job, buf, bb=None: calls.append(job.n_obs) or real(job, buf, bb))
    monkeypatch.setattr(defaults, "DEVICE_BLOCK_BYTES", want[0].nbytes * 7)
    print("case public out= memmap, traction order 1, 8 blocks", flush=True)
    mm = np.lib.format.open_memmap(tmp_path / "t.npy", mode="w+", dtype=np.float64, shape=want.shape)
    M.traction_matrix(obs, nrm, tris, mat, 0.1, order=1, out=mm)
    mm.flush()
    assert np.array_equal(np.load(tmp_path / "t.npy"), want)
    o32 = np.zeros(want.shape, np.float32)
    M.traction_matrix(obs, nrm, tris, mat, 0.1, order=1, out=o32)
    assert np.array_equal(o32, want.astype(np.float32))
    print("case public out= interaction order 1", flush=True)
    monkeypatch.setattr(defaults, "DEVICE_BLOCK_BYTES", wint[0].nbytes * 5)
    oi = np.zeros_like(wint)
    M.interaction_matrix(tris, mat, 0.1, order=1, out=oi)
    assert np.array_equal(oi, wint)
    assert calls == [50, 50, wint.shape[0]]


def test_source_major_layout(M):
    """The consumer layout (n_src, n_b, n_a, n_rec) copied on the device is
    bitwise the transpose of interaction_matrix, for 1, 2, 6 and 9
    components, tiles with ragged edges, a row range and the continuous
    layout; float32 out= and the device tensor too."""
    import torch
    from mhs_amd import consumer, workloads
    rng = np.random.default_rng(17)
    Pl = workloads.plane(10.0, 5.0, 9, 4, 65.0, -0.3)
    tris = workloads.orient_up(workloads.grid_mesh(Pl + rng.uniform(-0.05, 0.05, Pl.shape) * [1, 1, 0]))
    mat = M.Material(30.0, 25.0)
    cases = [dict(receiver="strike", source="strike"), dict(receiver="dip", source=("strike", "dip")),
             dict(receiver=("strike", "dip"), source=("strike", "dip", "tensile")), dict(),
             dict(order=1, rows=(5, 40)), dict(order=2, layout="continuous", shared=M.shared_nodes(tris))]
    for kw in cases:
        print(f"case source-major {kw.get('receiver')} {kw.get('source')} {kw.get('order', 0)}", flush=True)
        K = M.interaction_matrix(tris, mat, 0.1, **kw)
        S = consumer.interaction_source_major(tris, mat, 0.1, **kw)
        assert S.flags.c_contiguous and S.shape == (K.shape[1], K.shape[3], K.shape[2], K.shape[0])
        assert np.array_equal(S, np.transpose(K, (1, 3, 2, 0)))
    K = M.interaction_matrix(tris, mat, 0.1)
    out = np.zeros((K.shape[1], K.shape[3], K.shape[2], K.shape[0]), np.float32)
    consumer.interaction_source_major(tris, mat, 0.1, out=out)
    assert np.array_equal(out, np.transpose(K, (1, 3, 2, 0)).astype(np.float32))
    D = consumer.interaction_source_major(tris, mat, 0.1, device=True)
    assert isinstance(D, torch.Tensor) and D.is_cuda
    assert np.array_equal(D.cpu().numpy(), np.transpose(K, (1, 3, 2, 0)))


def test_identities_on_gpu(M, small):
    """total - eigenstress == elastic to rounding; interaction == stress
    contracted with the receiver plane and both frames (1e-13,
    verify_image_kernel.py:537-645); uniform P1/P2 nodal slip reproduces P0
    (1e-11, verify_image_kernel.py:439-534); continuous = summed columns of
    the discontinuous P2 matrix; negative self-diagonal."""
    from mhs_amd import tdcs
    g = small
    mat, tb, eb, ob = _mat(M, g), g["tb"], g["eb"], g["ob"]
    print("case identities", flush=True)
    # interaction = frame-resolved traction of the stress matrix
    pts = M.collocation_points(tb, 0)
    fr = tdcs.slip_frame(tb)
    S = M.stress_matrix(pts, tb, mat, eb)                       # (o, a, b, s, k)
    t = np.einsum("oabsk,ob->oask", S, fr[:, 2, :])
    want = np.einsum("oci,oisk,sdk->oscd", fr, t, fr)
    got = M.interaction_matrix(tb, mat, eb, obs_pts=pts)
    assert relmax(got, want) < 1e-13
    assert np.all(np.diagonal(M.interaction_matrix(tb, mat, eb, receiver="strike", source="strike")[:, :, 0, 0]) < 0)
    # continuous layout = discontinuous P2 columns summed into the shared nodes
    sn = M.shared_nodes(tb)
    Dd = M.disp_matrix(ob, tb, mat, eb, order=2)
    Dc = M.disp_matrix(ob, tb, mat, eb, order=2, layout="continuous")
    acc = np.zeros_like(Dc)
    for s in range(tb.shape[0]):
        for k in range(6):
            acc[:, :, sn.ids[s, k], :] += Dd[:, :, 6 * s + k, :]
    assert relmax(Dc, acc) < 1e-14


# The gate's own triangles (verify_image_kernel.py:89-101): buried and well
# clear; vertical and reaching z = 0, tilted so no component vanishes.
TRI_DEEP = np.array([[0.1, 0.0, -2.0], [1.0, 0.2, -2.1], [0.4, 0.9, -1.6]])
TRI_SURF = np.array([[0.0, 0.0, 0.0], [0.08, 1.0, 0.0], [0.03, 0.5, -1.0]])


def test_three_way_split_bitwise(M, small):
    """total - eigenstress == stress BITWISE, and parts=True equals the
    separate calls bitwise (verify_image_kernel.py:317-325: "an identity, not
    a tolerance"): the gate's two observers and triangles, the buried and the
    surface-reaching golden sets at every order, and the continuous layout."""
    g = small
    mat = M.Material(30.0, 45.0)
    cases = [("gate", np.array([[0.3, 0.4, -1.0], [0.0, 0.0, -0.05]]), np.array([TRI_SURF, TRI_DEEP]), 0.1, {}),
             ("buried p0", g["ob"], g["tb"], g["eb"], {}), ("buried p1", g["ob"], g["tb"], g["eb"], {"order": 1}),
             ("buried p2", g["ob"], g["tb"], g["eb"], {"order": 2}),
             ("surface p0", g["os"], g["ts"], float(g["es"]), {}),
             ("surface p2", g["os"], g["ts"], float(g["es"]), {"order": 2}),
             ("continuous", g["ob"], g["tb"], g["eb"], {"order": 2, "layout": "continuous"})]
    for name, obs, tris, eps, kw in cases:
        print(f"case split {name}", flush=True)
        sig = M.stress_matrix(obs, tris, mat, eps, **kw)
        tot = M.total_stress_matrix(obs, tris, mat, eps, **kw)
        star = M.eigenstress_matrix(obs, tris, mat, eps, **kw)
        assert np.abs(star).max() > 0.0
        assert np.array_equal(tot - star, sig), name
        s2, e2 = M.stress_matrix(obs, tris, mat, eps, parts=True, **kw)
        assert np.array_equal(s2, sig) and np.array_equal(e2, star), name
        t3, e3 = M.stress_matrix(obs, tris, mat, eps, parts=True, subtract_eigenstress=False, **kw)
        assert np.array_equal(t3, tot) and np.array_equal(e3, star), name


def test_partition_of_unity_every_entry_point(M):
    """Uniform nodal slip at P1/P2 reproduces P0 at every entry point, 1e-11
    (verify_image_kernel.py:439-534, the gate's triangles, seeded observers
    and receiver normal), the interaction with each receiver's K rows at its
    centroid; non-uniform nodal slip gives a different field (the tripwire
    against an assembler that ignores the nodal values)."""
    mat = M.Material(30.0, 45.0)
    tris = np.array([TRI_SURF, TRI_DEEP])
    eps, n = 0.1, 2
    rng = np.random.default_rng(8)
    obs = np.column_stack([rng.normal(size=7) * 1.5, rng.normal(size=7) * 1.5,
                           -np.abs(rng.normal(size=7)) * 1.5 - 0.3])
    from mhs_amd import tdcs
    nrm = np.repeat(tdcs.slip_frame(tris)[:1, 2, :], len(obs), axis=0)
    cen = M.collocation_points(tris, 0)
    cases = [("disp_matrix", lambda p: M.disp_matrix(obs, tris, mat, eps, order=p)),
             ("stress_matrix", lambda p: M.stress_matrix(obs, tris, mat, eps, order=p)),
             ("total_stress_matrix", lambda p: M.total_stress_matrix(obs, tris, mat, eps, order=p)),
             ("eigenstress_matrix", lambda p: M.eigenstress_matrix(obs, tris, mat, eps, order=p)),
             ("elastic_strain_matrix", lambda p: M.elastic_strain_matrix(obs, tris, mat, eps, order=p)),
             ("traction_matrix", lambda p: M.traction_matrix(obs, nrm, tris, mat, eps, order=p))]
    worst = 0.0
    for name, fn in cases:
        ref = fn(0)
        for p in (1, 2):
            K = (p + 1) * (p + 2) // 2
            got = fn(p)
            assert got.shape[-2] == n * K, name
            summed = got.reshape(got.shape[:-2] + (n, K, 3)).sum(axis=-2)
            print(f"case unity {name} P{p}", flush=True)
            worst = max(worst, relmax(summed, ref))
    # the interaction: receiver rows at the centroids (K per element), so the
    # P0 matrix is the row-and-column sum of each element's K x K block
    for rec, src in ((("strike", "dip", "normal"), ("strike", "dip", "tensile")), ("strike", "strike")):
        ref = M.interaction_matrix(tris, mat, eps, receiver=rec, source=src)
        for p in (1, 2):
            K = (p + 1) * (p + 2) // 2
            got = M.interaction_matrix(tris, mat, eps, order=p, receiver=rec, source=src,
                                       obs_pts=np.repeat(cen, K, axis=0))
            print(f"case unity interaction {rec} {src} P{p}", flush=True)
            summed = got[::K].reshape(n, n, K, *got.shape[2:]).sum(axis=2)
            worst = max(worst, relmax(summed, ref))
    assert worst < 1e-11, worst
    g1 = M.disp_matrix(obs, tris, mat, eps, order=1)
    uniform = np.zeros((n * 3, 3))
    uniform[:, 0] = 1.0
    spiked = np.zeros((n * 3, 3))
    spiked[0::3, 0] = 3.0
    du = np.einsum("oisk,sk->oi", g1, uniform)
    ds = np.einsum("oisk,sk->oi", g1, spiked)
    assert relmax(ds, du) > 1e-3


def test_float32_out_and_device_reuse(M, small):
    g = small
    mat, tb, eb, ob = _mat(M, g), g["tb"], g["eb"], g["ob"]
    out = np.zeros((ob.shape[0], 3, tb.shape[0], 3), dtype=np.float32)
    r = M.disp_matrix(ob, tb, mat, eb, out=out)
    assert r is out
    assert relmax(out, g["disp_b_p0"]) < 1e-6


def test_c3_full_matrix_one_call(M):
    """The whole n = 10k strike/strike matrix in one call (10^8 pairs in one
    set of launches) against the golden row blocks, so every row chunk and
    column tile of the launch geometry is exercised; then the mixed mode
    against FP64 over all 10^8 entries (stated bound 1e-6)."""
    from mhs_amd import workloads
    tris, eps = workloads.c3_interaction()
    g = np.load(GOLDEN / "c3_rows.npz")
    mat = M.Material(30.0, 30.0)
    print("case C3 full matrix fp64", flush=True)
    K = M.interaction_matrix(tris, mat, eps, receiver="strike", source="strike")
    for lo in (0, 5000, 5024, 9998):
        assert relmax(K[lo:lo + 2], g[f"rows_{lo}"]) < TOL, lo
    print("case C3 full matrix mixed", flush=True)
    with M.precision("mixed"):
        Km = M.interaction_matrix(tris, mat, eps, receiver="strike", source="strike")
    assert M.get_precision() == "fp64"
    err = relmax(Km, K)
    print(f"mixed relmax vs fp64 {err:.3e}", flush=True)
    assert err < 1e-6
    for lo in (0, 5000, 5024, 9998):
        assert relmax(Km[lo:lo + 2], g[f"rows_{lo}"]) < 1e-6, lo


def test_far_tiles_do_not_change_values(M):
    """The autotuned launch shape (column tiles per block of the far pass)
    only regroups pairs: every candidate gives bitwise the same matrix, and
    the tuner caches one of the candidates."""
    import torch
    from mhs_amd import api, runtime, workloads
    tris, eps = workloads.c3_interaction()
    for prec in ("fp64", "mixed"):
        print(f"case far tiles {prec}", flush=True)
        with M.precision(prec):
            job = api.plan(M.interaction_matrix, tris, M.Material(30.0, 30.0), eps, receiver="strike",
                           source="strike", rows=(4990, 5030))
        outs = []
        for ct in runtime.tile_candidates(prec == "mixed"):
            prep = runtime.Prepared(job)
            prep.far_tiles = ct
            with torch.cuda.stream(prep.stream):
                prep.launch()
            torch.cuda.synchronize()
            outs.append(prep.final.clone())
        assert all(torch.equal(outs[0], o) for o in outs[1:])
        prep = runtime.Prepared(job)
        ct = prep.tune()
        assert ct in runtime.tile_candidates(prec == "mixed")
        assert prep.dev.far_tiles[prep.tune_key()] == ct
