"""Block drivers with the reference's by_source / by_rows interface
(src/mhs/parallel.py:132-330): stitching, probe block, out= (including a
memmap), error messages, the reference's block planners, and the BLAS
environment guard -- on the CPU, with stand-in calls whose entries depend on
their own source (or row) only, as the matrices' entries do."""
import functools
import os

import numpy as np
import pytest

from mhs_amd import parallel as mp


def _per_source(obs, tris, *, K=3):
    """(n_obs, 3, n_src * K, 3): each source's K columns from its own vertices."""
    c = tris.mean(axis=1)                                   # (n, 3)
    d = obs[:, None, :] - c[None, :, :]                     # (n_obs, n, 3)
    r = np.sqrt((d * d).sum(-1) + 0.01)                     # (n_obs, n)
    k = np.arange(1, K + 1)
    base = r[:, :, None] * k[None, None, :]                 # (n_obs, n, K)
    out = base[:, None, :, :, None] * np.arange(1, 4)[None, :, None, None, None] \
        + tris[None, None, :, None, 0, :]                   # (n_obs, 3, n, K, 3)
    return out.reshape(obs.shape[0], 3, tris.shape[0] * K, 3)


def _interaction_like(tris, *, obs_tris):
    """(n_rec, n_src, 1, 1) with receivers fixed by obs_tris."""
    rc = obs_tris.mean(axis=1)
    c = tris.mean(axis=1)
    return np.linalg.norm(rc[:, None, :] - c[None, :, :], axis=-1)[:, :, None, None]


def _rows_call(M, rows):
    lo, hi = rows
    return M[lo:hi]


@pytest.fixture
def mesh():
    rng = np.random.default_rng(11)
    return rng.normal(size=(37, 3, 3)), rng.normal(size=(9, 3))


def test_by_source_stitches_columns(mesh):
    tris, obs = mesh
    call = functools.partial(_per_source, obs)
    want = call(tris)
    for bb in (None, 1, 500, 10 ** 9):
        got = mp.by_source(call, tris, workers=4, block_bytes=bb)
        assert got.shape == want.shape and np.array_equal(got, want)


def test_by_source_interaction_axis_and_keyword(mesh):
    tris, _ = mesh
    call = functools.partial(_interaction_like, obs_tris=tris)
    want = call(tris)
    got = mp.by_source(call, tris, source_axis=1, tris_kw="tris", block_bytes=100)
    assert np.array_equal(got, want)


def test_by_source_out_and_memmap(mesh, tmp_path):
    tris, obs = mesh
    call = functools.partial(_per_source, obs)
    want = call(tris)
    out = np.zeros_like(want)
    assert mp.by_source(call, tris, out=out, block_bytes=300) is out
    assert np.array_equal(out, want)
    mm = np.lib.format.open_memmap(tmp_path / "m.npy", mode="w+", dtype=want.dtype, shape=want.shape)
    mp.by_source(call, tris, out=mm, block_bytes=300)
    mm.flush()
    assert np.array_equal(np.load(tmp_path / "m.npy"), want)
    with pytest.raises(ValueError, match=r"out has shape \(1, 2\), expected"):
        mp.by_source(call, tris, out=np.zeros((1, 2)))


def test_by_rows(mesh):
    M = np.arange(5 * 300 * 4.0).reshape(300, 5, 4)
    call = functools.partial(_rows_call, M)
    for bb in (None, 64, 2000):
        assert np.array_equal(mp.by_rows(call, 300, block_bytes=bb), M)
    out = np.empty_like(M)
    assert mp.by_rows(call, 300, out=out) is out and np.array_equal(out, M)
    assert np.array_equal(mp.by_rows(call, 3), M[:3])
    with pytest.raises(ValueError, match="n_rows must be positive, got 0"):
        mp.by_rows(call, 0)
    with pytest.raises(ValueError, match="out has shape"):
        mp.by_rows(call, 300, out=np.empty((299, 5, 4)))


def test_planners_match_the_reference_rules():
    # at least one chunk per worker, more under a tight budget, none empty
    assert mp.plan_chunks(10, 4, 8) == [(0, 2), (2, 5), (5, 8), (8, 10)]
    ch = mp.plan_chunks(1000, 2, 1 << 20, budget=8 << 20)
    assert len(ch) == 250 and ch[0] == (0, 4) and ch[-1][1] == 1000
    assert mp.plan_chunks(0, 4, 8) == []
    assert mp.plan_chunks(3, 8, 8) == [(0, 1), (1, 2), (2, 3)]
    # row blocks never under ROWS_PER_BLOCK_MIN, one per worker when enough rows
    assert mp.plan_row_blocks(100, 4, 1 << 30) == [(0, 100)]
    rb = mp.plan_row_blocks(4096, 4, 8)
    assert len(rb) == 4 and all(b - a >= mp.ROWS_PER_BLOCK_MIN for a, b in rb)


def test_pinned_blas_restores_the_environment():
    key = mp.THREAD_ENV[0]
    before = os.environ.get(key)
    os.environ.pop(key, None)
    try:
        with mp.pinned_blas(3):
            assert all(os.environ[k] == "3" for k in mp.THREAD_ENV)
        assert key not in os.environ
    finally:
        if before is not None:
            os.environ[key] = before
