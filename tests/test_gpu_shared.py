"""The shared-node table on the device (``shared_nodes_device``, the C-ABI's
``mhs_shared_nodes``) against the host table of src/mhs/shared.py:175-273:
node ids, boundary and groups exactly, and coordinates / normals / weights
bitwise as well (the device sums each node's terms in the host's order with
the same IEEE operations, no FMA contraction), and every refusal with the
host's message."""
import numpy as np
import pytest

from mhs_amd import workloads
from mhs_amd.shared import shared_nodes, shared_nodes_device

pytestmark = pytest.mark.gpu


def _same(dev, host):
    assert dev.n_vertices == host.n_vertices and dev.tol == host.tol
    for name in ("ids", "boundary", "group"):
        a, b = getattr(dev, name), getattr(host, name)
        assert a.dtype == b.dtype and np.array_equal(a, b), name
    for name in ("coords", "normals", "weights"):
        a, b = getattr(dev, name), getattr(host, name)
        assert a.shape == b.shape, name
        assert np.array_equal(a, b), (name, float(np.max(np.abs(a - b))))


def _jittered_plane(nu, nv, seed):
    P = workloads.plane(60.0, 30.0, nu, nv, 50.0, -2.0)
    rng = np.random.default_rng(seed)
    inner = np.zeros(P.shape[:2], bool)
    inner[1:-1, 1:-1] = True
    P[inner] += rng.uniform(-0.05, 0.05, (int(inner.sum()), 3)) * (60.0 / nu)
    return workloads.orient_up(workloads.grid_mesh(P))


def test_network_with_groups():
    tris, groups, _ = workloads.c4_network()
    _same(shared_nodes_device(tris, groups=groups), shared_nodes(tris, groups=groups))
    # without groups the crossing faults weld where their vertices meet
    _same(shared_nodes_device(tris), shared_nodes(tris))


def test_jittered_planes_and_split_groups():
    for (nu, nv, seed) in ((1, 1, 0), (7, 3, 1), (160, 80, 2)):
        t = _jittered_plane(nu, nv, seed)
        _same(shared_nodes_device(t), shared_nodes(t))
    t = _jittered_plane(40, 20, 3)
    g = (np.arange(t.shape[0]) % 40 < 20).astype(int) * 7 - 3      # negative group ids too
    _same(shared_nodes_device(t, groups=g), shared_nodes(t, groups=g))
    tris, _ = workloads.c3_interaction()
    _same(shared_nodes_device(tris, tol_over_h=1e-6), shared_nodes(tris, tol_over_h=1e-6))


def _message(fn, *a, **k):
    with pytest.raises(ValueError) as e:
        fn(*a, **k)
    return str(e.value)


def test_refusals_match_the_host():
    A, B, C, D, E = (np.array(p, float) for p in
                     ([0, 0, -1], [1, 0, -1], [0, 1, -1], [1, 1, -1.5], [0.5, -1, -1.2]))
    cases = {
        "non-manifold": (np.array([[A, B, C], [B, A, D], [A, B, E]]), {}),
        "winding": (np.array([[A, B, C], [A, B, D]]), {}),
        "soup": (np.array([[A, B, C], [D + 5, E + 5, D + 6]]), {}),
        "weld tolerance": (np.array([[A, B, C], [B, D, C]]), {"tol_over_h": 3.0}),
        "collinear": (np.array([[A, B, 2 * B - A]]), {}),
        "zero-length edge": (np.array([[A, A, C]]), {}),
        "non-finite": (np.array([[A, B, C + np.nan]]), {}),
        "groups length": (np.array([[A, B, C], [B, D, C]]), {"groups": [0, 1, 2]}),
        "non-finite before groups": (np.array([[A, B, C + np.inf]]), {"groups": [0, 1]}),
    }
    for name, (t, kw) in cases.items():
        want = _message(shared_nodes, t, **kw)
        assert _message(shared_nodes_device, t, **kw) == want, name


def test_reference_golden_table(small):
    """The reference's own table of the golden mesh (ids, boundary exactly;
    coordinates, normals, weights to round-off)."""
    from conftest import relmax
    sn = shared_nodes_device(small["tb"])
    assert np.array_equal(sn.ids, small["sn_ids"])
    assert np.array_equal(sn.boundary, small["sn_boundary"])
    assert relmax(sn.coords, small["sn_coords"]) < 1e-15
    assert relmax(sn.normals, small["sn_normals"]) < 1e-15
    assert relmax(sn.weights, small["sn_weights"]) < 1e-15


def test_device_table_drives_the_continuous_layout():
    """The device-built table passed as ``shared=`` gives the continuous-layout
    matrices the host-built table gives, bitwise, on a grouped three-fault
    network (order 2, per-source eps)."""
    import mhs_amd as M
    tris, groups, eps = workloads.c4_network(target=300)
    mat = M.Material(30.0, 30.0)
    host, dev = shared_nodes(tris, groups=groups), shared_nodes_device(tris, groups=groups)
    kw = dict(order=2, layout="continuous", source=("strike", "dip"), receiver=("strike", "dip"))
    a = M.interaction_matrix(tris, mat, eps, shared=host, **kw)
    b = M.interaction_matrix(tris, mat, eps, shared=dev, **kw)
    assert a.shape == (host.n, host.n, 2, 2) and np.array_equal(a, b)
