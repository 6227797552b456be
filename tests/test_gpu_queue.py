"""The batched persistent path (mhs_assemble_batch): several launches of
different forms, orders, precisions, layouts and sizes in one batch, every
tiled pass drained from a device work queue by one persistent launch per
kernel instance, against each launch run alone through mhs_assemble --
bitwise; row ranges of one job over consecutive batches; and the serving
loop in queue mode, each result bitwise its alone result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mhs_amd
    return mhs_amd


@pytest.fixture(autouse=True, scope="module")
def _release_serving_arena():
    yield
    import torch
    from mhs_amd import serve
    serve.release_arenas()
    if torch.cuda.is_available():
        torch.cuda.empty_cache()


def _jobs(M):
    """(name, job) of every pass variant: displacement, stress (split and
    total), eigenstress, traction, the contracted interaction (one and two
    outputs, FP64 and mixed, grouped 3 x 3 in the mixed mode), the tensor
    interaction, the continuous layout; buried and surface-reaching."""
    from mhs_amd import api, workloads
    rng = np.random.default_rng(21)
    mat = M.Material(30.0, 25.0)
    buried = workloads.orient_up(workloads.grid_mesh(workloads.plane(6.0, 3.0, 6, 3, 55.0, -1.0)))
    surf = workloads.orient_up(workloads.grid_mesh(workloads.plane(6.0, 3.0, 5, 3, 70.0, 0.0)))
    obs = np.column_stack([rng.uniform(-5, 5, 90), rng.uniform(-5, 5, 90), -rng.uniform(0, 4, 90)])
    nrm = rng.normal(size=(90, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cases = [
        ("disp p0 surface", "fp64", M.disp_matrix, (obs, surf, mat, 0.1), {}),
        ("disp p2 buried", "mixed", M.disp_matrix, (obs, buried, mat, 0.1), {"order": 2}),
        ("stress p1 split", "fp64", M.stress_matrix, (obs[:40], surf, mat, 0.1), {"order": 1}),
        ("total stress p0", "mixed", M.total_stress_matrix, (obs[:30], buried, mat, 0.1), {}),
        ("eigenstress p1", "fp64", M.eigenstress_matrix, (obs[:30], surf, mat, 0.1), {"order": 1}),
        ("traction p1", "mixed", M.traction_matrix, (obs, nrm, surf, mat, 0.1), {"order": 1}),
        ("traction p2", "fp64", M.traction_matrix, (obs, nrm, buried, mat, 0.1), {"order": 2}),
        ("interaction ss p0 fp64", "fp64", M.interaction_matrix, (surf, mat, 0.1),
         {"receiver": "strike", "source": "strike"}),
        ("interaction ss p1 mixed", "mixed", M.interaction_matrix, (surf, mat, 0.1),
         {"order": 1, "receiver": "strike", "source": "strike"}),
        ("interaction 1x2 p2 mixed", "mixed", M.interaction_matrix, (buried, mat, 0.1),
         {"order": 2, "receiver": "dip", "source": ("strike", "dip")}),
        ("interaction 1x2 p0 fp64", "fp64", M.interaction_matrix, (surf, mat, 0.1),
         {"receiver": "normal", "source": ("strike", "dip")}),
        ("interaction 3x3 p2 mixed", "mixed", M.interaction_matrix, (surf, mat, 0.1), {"order": 2}),
        ("interaction 3x3 p1 fp64", "fp64", M.interaction_matrix, (buried, mat, 0.1), {"order": 1}),
        ("interaction continuous p2", "mixed", M.interaction_matrix, (surf, mat, 0.1),
         {"order": 2, "layout": "continuous", "source": ("strike", "dip")}),
        ("disp continuous p2", "fp64", M.disp_matrix, (obs[:50], surf, mat, 0.1),
         {"order": 2, "layout": "continuous"}),
    ]
    out = []
    for name, prec, fn, args, kw in cases:
        with M.precision(prec):
            out.append((name, api.plan(fn, *args, **kw)))
    return out


def _alone(job):
    import torch
    from mhs_amd import runtime
    prep = runtime.Prepared(job)
    with torch.cuda.stream(prep.stream):
        prep.launch()
    prep.stream.synchronize()
    prep.check_fold()
    return prep.final.clone(), (prep.final2.clone() if prep.final2 is not None else None)


def test_batch_equals_each_launch_alone(M):
    import torch
    from mhs_amd import runtime
    jobs = _jobs(M)
    want = [_alone(job) for _, job in jobs]
    stream = torch.cuda.Stream()
    preps = [runtime.Prepared(job, stream=stream) for _, job in jobs]
    print(f"case one batch of {len(preps)} launches", flush=True)
    b = runtime.Batch([(p, 0, p.rows) for p in preps], stream)
    b.done()
    for (name, _), p, (w, w2) in zip(jobs, preps, want):
        p.check_fold()
        if p.cont:
            # the continuous layout's scratch holds its rows: one batch per
            # row range (the range is the whole job here when it fits)
            assert p.rows >= p.job.n_obs, name
        d = float((p.final - w).abs().max())
        print(f"  {name}: max |batch - alone| {d:.3e}", flush=True)
        assert torch.equal(p.final, w), name
        if w2 is not None:
            assert torch.equal(p.final2, w2), name


def test_row_ranges_over_batches(M):
    """One job's rows cut into ranges over consecutive batches (each batch
    with other jobs' ranges) land bitwise its one-call matrix."""
    import torch
    from mhs_amd import runtime
    jobs = dict(_jobs(M))
    names = ["traction p1", "interaction 3x3 p2 mixed", "disp p0 surface"]
    want = [_alone(jobs[n])[0] for n in names]
    stream = torch.cuda.Stream()
    preps = [runtime.Prepared(jobs[n], stream=stream) for n in names]
    ranges = [p.chunks(max(1, p.n_cols * 7)) for p in preps]
    print("case row ranges over batches", flush=True)
    batches = []
    for r in range(max(len(c) for c in ranges)):
        items = [(p, *c[r]) for p, c in zip(preps, ranges) if r < len(c)]
        batches.append(runtime.Batch(items, stream))
    for b in batches:
        b.done()
    for n, p, w in zip(names, preps, want):
        p.check_fold()
        assert torch.equal(p.final, w), n


def test_queue_server_equals_alone(M):
    """The serving loop in queue mode (one stream, batches drained by
    persistent launches) on a mixed stream: every result bitwise alone."""
    import torch
    from mhs_amd import api, runtime, serve
    from test_gpu_serve import _requests
    reqs = _requests(M, serve)
    alone = []
    for r in reqs:
        with M.precision(r.precision):
            job = api.plan(r.fn, *r.args, **r.kwargs)
        out, _ = runtime.run(job)
        torch.cuda.synchronize()
        alone.append(out.clone())
    print("case queue server", flush=True)
    srv = serve.Server(slots=4, chunk_pairs=2000, queue=True)
    served = srv.run(reqs, keep=True)
    torch.cuda.synchronize()
    assert srv.batches > len(reqs)
    for r, a, s in zip(reqs, alone, served):
        assert s.error is None, s.error
        assert torch.equal(s.out, a), (r.fn.__name__, r.precision, float((s.out - a).abs().max()))
