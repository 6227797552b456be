"""Content cache key of the consumer layout (keyed like
studies/active_fractal/kernel.py:131-155): the same inputs give the same key,
and anything the matrix depends on -- vertex order included -- changes it."""
import numpy as np

import mhs_amd as M
from mhs_amd import consumer


def _mesh():
    rng = np.random.default_rng(3)
    return rng.normal(size=(12, 3, 3)) - [0, 0, 5.0]


def test_cache_key_is_stable_and_content_based():
    tris, mat = _mesh(), M.Material(30.0, 30.0)
    k = consumer.cache_key(tris, mat, 0.1, receiver="strike", source="strike")
    assert len(k) == 16 and int(k, 16) >= 0
    assert k == consumer.cache_key(tris.copy(), M.Material(30.0, 30.0), 0.1, source="strike", receiver="strike")
    variants = [
        consumer.cache_key(tris[::-1], mat, 0.1, receiver="strike", source="strike"),
        consumer.cache_key(tris[:, ::-1], mat, 0.1, receiver="strike", source="strike"),
        consumer.cache_key(tris, M.Material(30.0, 31.0), 0.1, receiver="strike", source="strike"),
        consumer.cache_key(tris, mat, 0.11, receiver="strike", source="strike"),
        consumer.cache_key(tris, mat, np.full(12, 0.1), receiver="strike", source="strike"),
        consumer.cache_key(tris, mat, 0.1, receiver="dip", source="strike"),
        consumer.cache_key(tris, mat, 0.1, receiver="strike", source="strike", order=1),
        consumer.cache_key(tris, mat, 0.1, receiver="strike", source="strike", layout_tag="row_major"),
    ]
    assert len(set(variants + [k])) == len(variants) + 1


def test_cache_key_precision_rule_and_normalised_keywords():
    """A matrix built in the mixed mode (FP32 far-field points) or under a
    fixed mixed rule is not the FP64 matrix: its key differs.  Keywords at
    their defaults, and 'normal' for 'tensile', are the same matrix: the
    same key."""
    from mhs_amd import runtime
    tris, mat = _mesh(), M.Material(30.0, 30.0)
    kw = dict(receiver="strike", source="strike")
    k64 = consumer.cache_key(tris, mat, 0.1, **kw)
    with M.precision("mixed"):
        kmix = consumer.cache_key(tris, mat, 0.1, **kw)
        old = runtime.mixed_rule
        runtime.mixed_rule = runtime.RULE_SAFE
        try:
            krule = consumer.cache_key(tris, mat, 0.1, **kw)
        finally:
            runtime.mixed_rule = old
    assert len({k64, kmix, krule}) == 3
    assert k64 == consumer.cache_key(tris, mat, 0.1, **kw)
    full = dict(receiver=("strike", "dip", "normal"), source=("strike", "dip", "tensile"), order=0,
                layout="discontinuous", subtract_eigenstress=True, fold=True)
    assert consumer.cache_key(tris, mat, 0.1) == consumer.cache_key(tris, mat, 0.1, **full)
    assert consumer.cache_key(tris, mat, 0.1) == consumer.cache_key(
        tris, mat, 0.1, receiver=("strike", "dip", "tensile"))
    assert consumer.cache_key(tris, mat, 0.1, **kw) == consumer.cache_key(
        tris, mat, 0.1, receiver=("strike",), source=["strike"])
    assert consumer.cache_key(tris, mat, 0.1) != consumer.cache_key(tris, mat, 0.1, fold=False)


def test_cache_key_per_source_eps_and_shared_nodes():
    tris, mat = _mesh(), M.Material(30.0, 30.0)
    e = np.linspace(0.05, 0.15, 12)
    e2 = e.copy()
    e2[3] += 1e-12
    assert consumer.cache_key(tris, mat, e) != consumer.cache_key(tris, mat, e2)
    from mhs_amd import workloads
    grid = workloads.orient_up(workloads.grid_mesh(workloads.plane(4.0, 2.0, 2, 1, 60.0, -1.0)))
    a = consumer.cache_key(grid, mat, 0.1, order=2, layout="continuous", shared=M.shared_nodes(grid))
    b = consumer.cache_key(grid, mat, 0.1, order=2, layout="continuous", shared=M.shared_nodes(grid.copy()))
    assert a == b
    assert a != consumer.cache_key(grid, mat, 0.1, order=2, layout="continuous")


def test_layout_and_stream_arguments_refused_before_device_work():
    """Bad out= (type, shape, dtype) and parts=True for streaming are
    refused from the validated plan alone, before anything reaches the
    device (this runs without a GPU)."""
    import pytest
    from mhs_amd import parallel
    from mhs_amd import workloads
    tris = workloads.orient_up(workloads.grid_mesh(workloads.plane(4.0, 2.0, 2, 1, 60.0, -1.0)))
    mat = M.Material(30.0, 30.0)
    kw = dict(receiver="strike", source=("strike", "dip"))
    n = tris.shape[0]
    with pytest.raises(ValueError, match=rf"out has shape \(1,\), expected \({n}, 2, 1, {n}\)"):
        consumer.interaction_source_major(tris, mat, 0.1, out=np.zeros(1), **kw)
    with pytest.raises(ValueError, match="float64 or float32"):
        consumer.interaction_source_major(tris, mat, 0.1, out=np.zeros((n, 2, 1, n), np.int64), **kw)
    with pytest.raises(TypeError, match="numpy.ndarray"):
        consumer.interaction_source_major(tris, mat, 0.1, out=[0.0], **kw)
    obs = np.array([[0.0, 1.0, -0.5], [1.0, 0.0, -1.0]])
    with pytest.raises(ValueError, match="parts=True"):
        parallel.stream_rows(M.stress_matrix, obs, tris, mat, 0.1, parts=True)
    with pytest.raises(ValueError, match=r"out has shape \(2, 3\), expected"):
        parallel.stream_rows(M.disp_matrix, obs, tris, mat, 0.1, out=np.zeros((2, 3)))
    with pytest.raises(ValueError, match="float64 or float32"):
        parallel.stream_rows(M.disp_matrix, obs, tris, mat, 0.1, out=np.zeros((2, 3, n, 3), np.float16))
    # the reference's own validation comes first, with its messages
    with pytest.raises(ValueError, match="z"):
        parallel.stream_rows(M.disp_matrix, obs * [1, 1, -1], tris, mat, 0.1)
