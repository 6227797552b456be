"""A prepared assembly captured as one HIP graph: every replay, with eager
work and synchronisations between replays, equals the eager result bit for
bit (contracted interaction with the fold, and an uncontracted traction)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mhs_amd as M
    from mhs_amd import api, runtime, workloads
    tris, eps = workloads.c3_interaction(nx=60, nz=10)
    mat = M.Material(30.0, 30.0)
    rng = np.random.default_rng(3)
    obs = np.column_stack([rng.uniform(-20, 20, 50), rng.uniform(-2, 2, 50), -rng.uniform(0, 5, 50)])
    nrm = rng.normal(size=(50, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cases = [("interaction mixed", "mixed", (M.interaction_matrix, (tris, mat, eps),
                                             {"receiver": "strike", "source": "strike"})),
             ("traction fp64", "fp64", (M.traction_matrix, (obs, nrm, tris, mat, eps), {"order": 1}))]
    for name, prec, (fn, args, kw) in cases:
        print(f"case graph {name}", flush=True)
        with M.precision(prec):
            job = api.plan(fn, *args, **kw)
        eager = runtime.Prepared(job, stream=torch.cuda.Stream())
        with torch.cuda.stream(eager.stream):
            eager.launch()
        torch.cuda.synchronize()
        want = eager.final.clone()
        prep = runtime.Prepared(job, stream=torch.cuda.Stream())
        g = prep.graph()
        other = runtime.Prepared(job, stream=torch.cuda.Stream())
        for r in range(3):
            with torch.cuda.stream(prep.stream):
                prep.final.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(prep.final, want), (name, r)
            with torch.cuda.stream(other.stream):      # eager work between replays
                other.launch()
            torch.cuda.synchronize()
            assert torch.equal(other.final, want)
        prep.check_fold()
