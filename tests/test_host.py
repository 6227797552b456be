"""Host-side drop-in boundary: signatures, validation and refusals, frames,
shared nodes, collocation points and byte arithmetic -- all before any GPU
work (tests/gates/mhs/verify_import_hygiene.py:159-271,
verify_shared_nodes.py:294-328)."""
import inspect

import numpy as np
import pytest

import mhs_amd as M
from conftest import relmax

EXPECTED = {
    "disp_matrix": ["obs_pts", "tris", "material", "eps", "order", "layout", "shared", "out", "fold"],
    "stress_matrix": ["obs_pts", "tris", "material", "eps", "subtract_eigenstress", "parts", "order",
                      "layout", "shared", "out", "fold"],
    "total_stress_matrix": ["obs_pts", "tris", "material", "eps", "order", "layout", "shared", "out", "fold"],
    "eigenstress_matrix": ["obs_pts", "tris", "material", "eps", "order", "layout", "shared", "out", "fold"],
    "elastic_strain_matrix": ["obs_pts", "tris", "material", "eps", "subtract_eigenstress", "order",
                              "layout", "shared", "out", "fold"],
    "traction_matrix": ["obs_pts", "obs_normals", "tris", "material", "eps", "subtract_eigenstress", "order",
                        "layout", "shared", "out", "fold"],
    "interaction_matrix": ["tris", "material", "eps", "receiver", "source", "obs_tris", "obs_pts", "order",
                           "shrink", "subtract_eigenstress", "layout", "shared", "rows", "out", "fold"],
    "shared_nodes": ["tris", "groups", "tol_over_h"],
}

MAT = M.Material(mu=30.0, lam=45.0)
TRI = np.array([[[0.0, 0.0, -2.0], [1.0, 0.0, -2.0], [0.0, 1.0, -2.0]]])
OBS = np.array([[0.3, 0.2, -1.0]])


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_signatures(name):
    fn = getattr(M, name)
    assert list(inspect.signature(fn).parameters) == EXPECTED[name]


def test_defaults():
    for name in ("disp_matrix", "stress_matrix", "traction_matrix", "interaction_matrix"):
        params = inspect.signature(getattr(M, name)).parameters
        assert params["order"].default == 0
        assert params["layout"].default == "discontinuous"
    assert inspect.signature(M.stress_matrix).parameters["subtract_eigenstress"].default is True
    assert not hasattr(M, "strain_matrix")


@pytest.mark.parametrize("call,exc", [
    (lambda: M.disp_matrix(np.array([[0.0, 0.0, 1.0]]), TRI, MAT, 0.1), ValueError),
    (lambda: M.disp_matrix(OBS, TRI + np.array([0.0, 0.0, 3.0]), MAT, 0.1), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.0), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, -1.0), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, 0.25, 0.1), TypeError),
    (lambda: M.disp_matrix(OBS, TRI[0], MAT, 0.1), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.1, order=3), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.1, layout="weird"), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.1, layout="continuous"), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, np.array([0.1, 0.2])), ValueError),
    (lambda: M.disp_matrix(np.zeros((0, 3)), TRI, MAT, 0.1), ValueError),
    (lambda: M.disp_matrix(OBS * np.nan, TRI, MAT, 0.1), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.1, out=[1]), TypeError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.1, out=np.zeros((1, 3, 1, 2))), ValueError),
    (lambda: M.disp_matrix(OBS, TRI, MAT, 0.1, out=np.zeros((1, 3, 1, 3), dtype=np.int64)), ValueError),
    (lambda: M.stress_matrix(OBS, TRI, MAT, 0.1, parts=True, out=np.zeros((1, 3, 3, 1, 3))), ValueError),
    (lambda: M.traction_matrix(OBS, np.zeros(3), TRI, MAT, 0.1), ValueError),
    (lambda: M.traction_matrix(OBS, np.ones((2, 3)), TRI, MAT, 0.1), ValueError),
    (lambda: M.interaction_matrix(TRI, MAT, 0.1, rows=(0, 5)), ValueError),
    (lambda: M.interaction_matrix(TRI, MAT, 0.1, receiver="sideways"), ValueError),
    (lambda: M.interaction_matrix(TRI, MAT, 0.1, obs_pts=np.zeros((2, 3)) - 1.0), ValueError),
    (lambda: M.Material.from_mu_nu(30.0, 0.5), ValueError),
    (lambda: M.Material(0.0, 1.0), ValueError),
    (lambda: M.Material(1.0, -1.0), ValueError),
])
def test_refusals_before_any_gpu_work(call, exc):
    with pytest.raises(exc):
        call()


def test_ceiling_refuses_with_arithmetic():
    big_tris = np.repeat(TRI, 2000, axis=0)
    big_obs = np.repeat(OBS, 30_000, axis=0)
    with pytest.raises(MemoryError, match="per obs/source pair: 216 B"):
        M.stress_matrix(big_obs, big_tris, MAT, 0.1)


def test_byte_arithmetic():
    assert M.chunking.pair_bytes("stress") == 216
    assert M.chunking.pair_bytes("stress", parts=True) == 432
    assert M.chunking.pair_bytes("disp") == 72
    assert M.chunking.pair_bytes((1, 1)) == 8
    plan = M.chunking.chunk_plan(10_000, 10_000, "stress", ceiling=8 << 30)
    assert plan.bytes_total == 216 * 10**8
    assert plan.n_chunks * plan.obs_per_chunk >= 10_000


def test_material_roundtrip():
    m = M.Material.from_mu_nu(30.0, 0.25)
    assert abs(m.lam - 30.0) < 1e-12
    assert abs(m.nu_if_you_must - 0.25) < 1e-15


def test_frames_and_shared_nodes_match_reference(small):
    g = small
    assert relmax(M.tdcs.slip_frame(g["tb"]), g["frame_b"]) < 1e-15
    assert relmax(M.tdcs.slip_frame(g["ts"]), g["frame_s"]) < 1e-15
    sn = M.shared_nodes(g["tb"])
    assert np.array_equal(sn.ids, g["sn_ids"])
    assert relmax(sn.coords, g["sn_coords"]) < 1e-15
    assert relmax(sn.normals, g["sn_normals"]) < 1e-15
    assert relmax(sn.weights, g["sn_weights"]) < 1e-15
    assert np.array_equal(sn.boundary, g["sn_boundary"])


def test_shared_nodes_refusals():
    a = np.array([[0, 0, -1.0], [1, 0, -1.0], [0, 1, -1.0]])
    b = np.array([[1, 0, -1.0], [1, 1, -1.0], [0, 1, -1.0]])
    far = a + np.array([5.0, 0, 0])
    with pytest.raises(ValueError, match="nothing is shared"):
        M.shared_nodes(np.stack([a, far]))
    with pytest.raises(ValueError, match="winding"):
        M.shared_nodes(np.stack([a, b[[0, 2, 1]]]))
    c = np.array([[1, 0, -1.0], [0, 1, -1.0], [0.5, 0.5, -2.0]])
    with pytest.raises(ValueError, match="non-manifold"):
        M.shared_nodes(np.stack([a, b, c]))
    sn = M.shared_nodes(np.stack([a, b]))
    assert sn.n == 4 + 5 and sn.ids.shape == (2, 6)


def test_shared_nodes_device_refuses_shapes_before_device_work():
    """A malformed array is refused by the device table before any GPU call,
    with the host table's message (it runs here without one)."""
    for bad in (np.zeros((2, 3)), np.zeros((0, 3, 3)), np.zeros((2, 3, 2))):
        with pytest.raises(ValueError, match="tris must have shape") as host:
            M.shared_nodes(bad)
        with pytest.raises(ValueError, match="tris must have shape") as dev:
            M.shared_nodes_device(bad)
        assert str(dev.value) == str(host.value)


def test_collocation_points():
    tri = TRI[0]
    p0 = M.collocation_points(TRI, 0)
    assert np.allclose(p0, tri.mean(axis=0))
    p1 = M.collocation_points(TRI, 1, shrink=0.5)
    assert np.allclose(p1[0], 0.5 * tri[0] + 0.5 * tri.mean(axis=0))
    p2 = M.collocation_points(TRI, 2, shrink=0.0)
    assert np.allclose(p2[3], 0.5 * (tri[0] + tri[1]))


def test_component_index():
    assert list(M.tdcs.component_index(("strike", "normal"))) == [0, 2]
    with pytest.raises(ValueError):
        M.tdcs.component_index(())


def test_precision_setting():
    import mhs_amd
    assert mhs_amd.get_precision() == "fp64"
    with mhs_amd.precision("mixed"):
        assert mhs_amd.get_precision() == "mixed"
        with mhs_amd.precision("fp64"):
            assert mhs_amd.get_precision() == "fp64"
        assert mhs_amd.get_precision() == "mixed"
    assert mhs_amd.get_precision() == "fp64"
    with pytest.raises(ValueError):
        mhs_amd.set_precision("fp16")
    with pytest.raises(ValueError):
        mhs_amd.precision("half")


def test_flattened_gauss_rules_match_the_nested_rule():
    """The flattened collapsed-Gauss table the far pass reads equals the
    oracle's nested rule (moments.py:427-498) point for point."""
    from mhs_amd import geometry
    from oracle.geometry import gauss_triangle
    t = geometry.gauss_triangle_tables().reshape(3, -1)
    for n in (1, 2, 8, 12, 37, 192):
        off = (n - 1) * n * (2 * n - 1) // 6
        x1, x2, w = gauss_triangle(n)
        assert np.array_equal(t[0, off:off + n * n], x1)
        assert np.array_equal(t[1, off:off + n * n], x2)
        assert np.allclose(t[2, off:off + n * n], w, rtol=1e-15, atol=0)
        assert abs(t[2, off:off + n * n].sum() - 0.5) < 1e-14


def test_fold_work_bound_covers_the_band():
    """The host bound that sizes the fold work buffer is the band count of
    the branch histogram (fold.py:84-118) plus its margin."""
    from mhs_amd import flops, geometry, runtime, workloads
    tris, eps = workloads.c3_interaction(nx=40, nz=6)
    obs = M.collocation_points(tris, 0)
    recs = geometry.source_records(tris, np.full(tris.shape[0], eps))
    strad = np.flatnonzero(recs[:, geometry.S_STRADDLE] != 0.0)
    h = flops.histogram(obs, tris, eps)
    assert h["fold_pairs"] > 0
    assert runtime.fold_pairs_bound(obs, recs, strad) == h["fold_pairs"] + 64


def test_rule_tuner_prefers_the_most_accurate_of_near_equal_rules():
    """The mixed rule tuner's choice (runtime.pick_rule): candidates beyond
    the error bound never win; among those within it, sample times within
    RULE_TIE of the fastest are tied and the smallest error wins; a clearly
    faster candidate wins over a more accurate one; nothing within the bound
    gives the fallback.  The table is a measured one (interaction n=10k,
    sample errors and ms of four rules)."""
    from mhs_amd import runtime as R
    measured = {
        "5 L first band": (1.82e-8, 7.651),
        "3 L first band": (9.64e-8, 7.620),
        "3.5 L first band": (3.25e-8, 7.652),
        "3 L switch": (7.22e-9, 7.878),
        "1 L fold switch": (4.6e-7, 7.10),
    }
    assert R.pick_rule(measured, bound=1e-7, tie=0.01) == "5 L first band"
    # a wider tie reaches the 3 L switch (3.4% slower, most accurate)
    assert R.pick_rule(measured, bound=1e-7, tie=0.05) == "3 L switch"
    # no tie: the fastest within the bound
    assert R.pick_rule(measured, bound=1e-7, tie=0.0) == "3 L first band"
    # the bound excludes the fastest of all
    assert R.pick_rule(measured, bound=1e-6, tie=0.0) == "1 L fold switch"
    assert R.pick_rule({"a": (2e-7, 1.0)}, bound=1e-7, fallback="safe") == "safe"
    # equal errors: the faster, then the earlier
    assert R.pick_rule({"a": (1e-8, 1.005), "b": (1e-8, 1.0), "c": (1e-8, 1.0)}, tie=0.01) == "b"
