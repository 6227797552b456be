"""mhs_assemble_static, the allocation-free variant of mhs_assemble: the
fold's work buffers come from the library's static device scratch.  Through
the C-ABI it must give bitwise the matrix of mhs_assemble on the same
inputs, on two calls in a row (the second smaller, so the scratch holds the
first call's leftovers), for the contracted interaction and the stress form
of a fault reaching the surface (folded pairs present); a launch whose
folded pairs exceed the scratch is refused with a message.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mhs_amd
    from mhs_amd import _lib
    return mhs_amd, _lib.load(), torch


def _mesh(nx, nz, seed):
    from mhs_amd import workloads
    rng = np.random.default_rng(seed)
    P = workloads.plane(float(nx), 0.5 * nz, nx, nz, 75.0, 0.0)
    P = P + rng.uniform(-0.02, 0.02, P.shape) * [1, 1, 0]
    tris = workloads.orient_up(workloads.grid_mesh(P))
    tris[:, :, 2] = np.minimum(tris[:, :, 2], 0.0)
    return tris


def _pair(M, lib, torch, job):
    """(regular, static) outputs of one launch over every row of ``job``."""
    from mhs_amd import _lib, geometry, runtime
    prep = runtime.Prepared(job)
    prep.far_tiles = 16
    lo, hi = 0, job.n_obs
    assert prep.n_strad > 0
    with torch.cuda.stream(prep.stream):
        L = prep._launch_struct(lo, hi, C.c_void_p(prep.final.data_ptr()),
                                C.c_void_p(prep.final2.data_ptr()) if prep.final2 is not None else None, 0)
        _lib.check(lib.mhs_assemble(C.byref(L), prep.handle), lib)
        out = torch.full_like(prep.final, float("nan"))
        out2 = torch.full_like(prep.final2, float("nan")) if prep.final2 is not None else None
        err = torch.full((1,), 7, dtype=torch.int32, device=prep.final.device)
        recs = geometry.source_records(job.tris, job.eps)
        strad = np.flatnonzero(recs[:, geometry.S_STRADDLE] != 0.0)
        S = prep._launch_struct(lo, hi, C.c_void_p(out.data_ptr()),
                                C.c_void_p(out2.data_ptr()) if out2 is not None else None, 0)
        S.fold_list = S.fold_count = S.fold_work = None
        S.fold_cap = 0
        S.fold_work_items = runtime.fold_pairs_bound(job.obs[lo:hi], recs, strad)
        S.fold_err = C.c_void_p(err.data_ptr())
        assert 0 < S.fold_work_items <= 3200
        _lib.check(lib.mhs_assemble_static(C.byref(S), prep.handle), lib)
    torch.cuda.synchronize()
    prep.check_fold()
    assert int(err.item()) == 0
    return (prep.final, out), (prep.final2, out2), prep


def _job(M, kind, tris):
    from mhs_amd import api
    mat = M.Material(30.0, 30.0)
    if kind == "interaction":
        return api.plan(M.interaction_matrix, tris, mat, 0.1, receiver="strike", source="strike")
    obs = M.collocation_points(tris, 0)
    return api.plan(M.stress_matrix, obs, tris, mat, 0.1)


@pytest.mark.parametrize("kind", ["interaction", "stress"])
def test_static_equals_allocating_entry_point(env, kind):
    M, lib, torch = env
    for step, (nx, nz) in enumerate(((16, 6), (10, 4))):      # the second call is smaller
        print(f"case static {kind} call {step + 1}: {2 * nx * nz} triangles", flush=True)
        job = _job(M, kind, _mesh(nx, nz, 40 + step))
        (a, b), (a2, b2), _ = _pair(M, lib, torch, job)
        assert torch.equal(a, b), (kind, step)
        if a2 is not None:
            assert torch.equal(a2, b2), (kind, step)


def test_static_refuses_too_many_folded_pairs(env):
    M, lib, torch = env
    from mhs_amd import _lib, runtime
    job = _job(M, "interaction", _mesh(10, 4, 7))
    prep = runtime.Prepared(job)
    prep.far_tiles = 16
    err = torch.zeros(1, dtype=torch.int32, device=prep.final.device)
    S = prep._launch_struct(0, job.n_obs, C.c_void_p(prep.final.data_ptr()), None, 0)
    S.fold_err = C.c_void_p(err.data_ptr())
    S.fold_work_items = 3201
    assert lib.mhs_assemble_static(C.byref(S), prep.handle) != 0
    assert b"static scratch" in lib.mhs_last_error()
    S.fold_err = None
    S.fold_work_items = 10
    assert lib.mhs_assemble_static(C.byref(S), prep.handle) != 0
    assert b"fold_err" in lib.mhs_last_error()
    torch.cuda.synchronize()
