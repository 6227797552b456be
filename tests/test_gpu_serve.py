"""Served stream on the GPU: each request's result is the same alone and in
the mix (several requests in flight on their own streams, cut into small row
chunks), in both precisions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(M, serve):
    from mhs_amd import workloads
    rng = np.random.default_rng(5)
    mat = M.Material(30.0, 25.0)
    out = []
    for i, (kind, order, prec) in enumerate([("disp", 0, "fp64"), ("interaction", 1, "mixed"),
                                             ("traction", 2, "fp64"), ("interaction", 0, "fp64"),
                                             ("disp", 1, "mixed"), ("traction", 0, "mixed")]):
        r = workloads.c5_request(rng, kind_u={"disp": 0.1, "traction": 0.5, "interaction": 0.9}[kind])
        tris = r["tris"][:60]
        if kind == "disp":
            g = np.linspace(-8, 8, 9)
            X, Y = np.meshgrid(g, g)
            obs = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])
            req = serve.Request(M.disp_matrix, (obs, tris, mat, 0.1), {"order": order})
        elif kind == "traction":
            obs = np.column_stack([rng.uniform(-8, 8, 70), rng.uniform(-8, 8, 70), -rng.uniform(0, 6, 70)])
            nrm = rng.normal(size=(70, 3))
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            req = serve.Request(M.traction_matrix, (obs, nrm, tris, mat, 0.1), {"order": order})
        else:
            req = serve.Request(M.interaction_matrix, (tris, mat, 0.1),
                                {"order": order, "receiver": "strike", "source": "strike"})
        req.precision = prec
        out.append(req)
    return out


def test_served_equals_alone():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mhs_amd as M
    from mhs_amd import api, runtime, serve
    reqs = _requests(M, serve)
    alone = []
    for r in reqs:
        with M.precision(r.precision):
            job = api.plan(r.fn, *r.args, **r.kwargs)
        out, _ = runtime.run(job)
        torch.cuda.synchronize()
        alone.append(out.clone())
    print("case served mix", flush=True)
    served = serve.Server(slots=3, chunk_pairs=2000).run(reqs, keep=True)
    torch.cuda.synchronize()
    for r, a, s in zip(reqs, alone, served):
        assert s.error is None, s.error
        assert s.out.shape == a.shape
        assert torch.equal(s.out, a), (r.fn.__name__, r.precision, float((s.out - a).abs().max()))
    summ = serve.summary(served)
    assert summ["completed"] == len(reqs)
