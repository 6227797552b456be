"""Host logic of the serving loop (mhs_amd/serve.py) with a CPU stand-in for
the device: first-come first-served admission by predicted bytes, the slot
limit, refusal of a request larger than the budget, validation errors as
results, chunked issue, and the latency summary."""
import time

import numpy as np
import pytest

import mhs_amd as M
from mhs_amd import runtime, serve, workloads


class _Event:
    def __init__(self, t_done):
        self.t_done = t_done

    def query(self):
        return time.perf_counter() >= self.t_done


class _Handle:
    def __init__(self, job):
        self.job = job

    def chunks(self, max_pairs):
        step = max(1, max_pairs // (self.job.n_src * self.job.K))
        return [(lo, min(self.job.n_obs, lo + step)) for lo in range(0, self.job.n_obs, step)]


class FakeBackend:
    """Chunks complete 1 ms after issue; records the slots in use."""

    def __init__(self):
        self.live = set()
        self.max_live = 0
        self.issued = []

    def default_budget(self):
        return 1 << 40

    def prepare(self, job, slot):
        assert slot not in self.live
        self.live.add(slot)
        self.max_live = max(self.max_live, len(self.live))
        h = _Handle(job)
        h.slot = slot
        h.serial = self.n_prepared = getattr(self, "n_prepared", 0) + 1
        return h

    def issue(self, h, lo, hi):
        self.issued.append((h.serial, lo, hi))
        return _Event(time.perf_counter() + 1e-3)

    def finish(self, h):
        self.live.discard(h.slot)
        return ("out", h.job.kind), None


def _disp_request(n_obs, arrival):
    obs, tris, eps = workloads.c1_disp()
    return serve.Request(M.disp_matrix, (obs[:n_obs], tris[:16], M.Material(30.0, 30.0), eps),
                         arrival=arrival)


def test_all_complete_with_slot_limit_and_chunks():
    be = FakeBackend()
    reqs = [_disp_request(64 + 8 * i, 0.0) for i in range(7)]
    srv = serve.Server(slots=3, chunk_pairs=16 * 20, backend=be)
    out = srv.run(reqs)
    assert all(s.error is None for s in out)
    assert be.max_live <= 3
    assert all(s.out == ("out", "disp") for s in out)
    # every request's rows are issued exactly once, in chunks of <= 20 rows
    rows = {}
    for h, lo, hi in be.issued:
        assert hi - lo <= 20
        rows.setdefault(h, []).append((lo, hi))
    assert sorted(sum(hi - lo for lo, hi in v) for v in rows.values()) == sorted(64 + 8 * i for i in range(7))
    assert all(s.latency >= 0 and s.admitted >= s.arrival - 1e-9 for s in out)
    summ = serve.summary(out)
    assert summ["completed"] == 7 and summ["p99_s"] >= summ["p50_s"] > 0


def test_admission_is_fifo_by_bytes():
    be = FakeBackend()
    reqs = [_disp_request(200, 0.0), _disp_request(10, 0.0), _disp_request(200, 0.0)]
    big = runtime.job_bytes(M.api.plan(reqs[0].fn, *reqs[0].args))
    small = runtime.job_bytes(M.api.plan(reqs[1].fn, *reqs[1].args))
    # room for one big request at a time: the small one waits behind the first big one
    srv = serve.Server(slots=4, budget_bytes=big + small // 2, backend=be)
    out = srv.run(reqs)
    assert all(s.error is None for s in out)
    assert out[1].admitted >= out[0].done - 1e-3
    assert out[2].admitted >= out[1].admitted


def test_oversized_and_invalid_requests_are_results():
    be = FakeBackend()
    ok = _disp_request(10, 0.0)
    bad = serve.Request(M.disp_matrix, (np.zeros((4, 3)) + [0, 0, 1.0], ok.args[1], ok.args[2], 0.1))
    srv = serve.Server(slots=2, budget_bytes=1000, backend=be)
    out = srv.run([ok, bad])
    assert out[0].error.startswith("MemoryError")
    assert out[1].error.startswith("ValueError")
    assert serve.summary(out)["refused"] == 2


def test_realtime_arrivals_are_respected():
    be = FakeBackend()
    reqs = [_disp_request(16, 0.0), _disp_request(16, 0.05)]
    out = serve.Server(slots=2, backend=be).run(reqs)
    assert out[1].admitted >= 0.05 - 1e-3


def test_c5_stream_shapes():
    reqs = serve.c5_requests(n_req=40, rate=10.0, seed=7)
    kinds = [r.tag.split()[0] for r in reqs]
    assert set(kinds) <= {"disp", "traction", "interaction"}
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    for r in reqs:
        job = M.api.plan_device(r.fn, *r.args, **r.kwargs)
        if job.kind == "interaction":
            # uncut meshes, the reference's default 3 x 3 components
            assert job.n_obs == job.n_src * job.K and (job.n_a, job.n_b) == (3, 3)
        assert runtime.job_bytes(job) > 0


def test_device_plans_skip_the_host_ceiling_only():
    """A served (device-resident) request is admitted against the device
    budget: plan_device skips the host byte ceiling that plan (and the
    public call) keeps, with the reference's MemoryError."""
    from mhs_amd import defaults
    P = workloads.plane(40.0, 20.0, 80, 40, 60.0, -1.0)
    tris = workloads.orient_up(workloads.grid_mesh(P))        # 6,400 triangles
    mat = M.Material(30.0, 30.0)
    kw = {"order": 2}                                       # (38400, 38400, 3, 3): 106 GB
    with pytest.raises(MemoryError):
        M.api.plan(M.interaction_matrix, tris, mat, 0.1, **kw)
    job = M.api.plan_device(M.interaction_matrix, tris, mat, 0.1, **kw)
    assert runtime.job_bytes(job) > defaults.MATRIX_MAX_BYTES
    with pytest.raises(MemoryError):                        # the flag does not leak into later calls
        M.api.plan(M.interaction_matrix, tris, mat, 0.1, **kw)


def test_cache_is_trimmed_as_requests_finish():
    """The loop gives the backend a chance to return cached device memory
    (headroom for the runtime's scratch) whenever requests complete."""
    class Trimming(FakeBackend):
        trims = 0

        def trim(self):
            self.trims += 1

    be = Trimming()
    reqs = [_disp_request(64, 0.0) for _ in range(6)]
    done = serve.Server(slots=2, backend=be, chunk_pairs=1 << 14).run(reqs, keep=False, realtime=False)
    assert all(s.error is None for s in done)
    assert 1 <= be.trims <= len(reqs)


def test_results_are_released_when_not_kept():
    """keep=False: no reference to a finished request's result survives its
    completion (its bytes are returned to the budget then)."""
    import gc
    import weakref

    class Result:
        pass

    class Releasing(FakeBackend):
        def __init__(self):
            super().__init__()
            self.refs = []

        def finish(self, h):
            super().finish(h)
            gc.collect()
            self.alive = max(getattr(self, "alive", 0), sum(r() is not None for r in self.refs))
            r = Result()
            self.refs.append(weakref.ref(r))
            return r, None

    be = Releasing()
    reqs = [_disp_request(32, 0.0) for _ in range(4)]
    done = serve.Server(slots=2, backend=be).run(reqs, keep=False, realtime=False)
    gc.collect()
    assert all(s.error is None and s.out is None for s in done)
    assert len(be.refs) == 4 and all(r() is None for r in be.refs)
    assert be.alive == 0          # none survived while the loop ran either


@pytest.mark.parametrize("bypass_s,small", [(0.0, False), (60.0, False), (0.0, True)])
def test_requests_that_fit_pass_a_blocked_one_for_a_while(bypass_s, small):
    """First fit within ``bypass_s``: a request behind a large one that does
    not fit yet is admitted past it; after that only requests below
    ``small_bytes`` pass; with both 0 the queue is strictly first come,
    first served."""
    be = FakeBackend()
    reqs = [_disp_request(200, 0.0), _disp_request(200, 0.0), _disp_request(10, 0.0)]
    big = runtime.job_bytes(M.api.plan(reqs[0].fn, *reqs[0].args))
    tail = runtime.job_bytes(M.api.plan(reqs[2].fn, *reqs[2].args))
    srv = serve.Server(slots=4, budget_bytes=big + tail, backend=be, bypass_s=bypass_s,
                       small_bytes=tail + 1 if small else 0)
    out = srv.run(reqs, realtime=False)
    assert all(s.error is None for s in out)
    if bypass_s > 0 or small:
        assert out[2].admitted < out[1].admitted
    else:
        assert out[2].admitted >= out[1].admitted


def test_arena_ends_and_coalescing():
    a = serve.Arena(10_000, small=1000, align=100)
    big = a.alloc(2950)                 # rounds up to 3000, from the low end
    small = a.alloc(150)                # 200, from the high end
    mid = a.alloc(5000)
    assert (big, small, mid) == (0, 9800, 3000)
    assert a.free == [(8000, 1800)] and a.alloc(2000) is None
    a.release(mid, 5000)
    a.release(big, 2950)
    assert a.free == [(0, 9800)]        # coalesced with both neighbours
    with pytest.raises(ValueError):
        a.release(0, 100)               # not allocated
    a.release(small, 150)
    assert a.free == [(0, 10_000)] and a.largest == 10_000


def test_arena_random_alloc_release_keeps_invariants():
    rng = np.random.default_rng(3)
    a = serve.Arena(1 << 20, small=1 << 14, align=256)
    live = []
    for _ in range(2000):
        if live and rng.uniform() < 0.45:
            off, n = live.pop(int(rng.integers(len(live))))
            a.release(off, n)
        else:
            n = int(np.exp(rng.uniform(np.log(10), np.log(1 << 17))))
            off = a.alloc(n)
            if off is not None:
                live.append((off, n))
        spans = sorted(a.free + [(o, a._round(n)) for o, n in live])
        assert all(o1 + s1 == o2 for (o1, s1), (o2, _) in zip(spans, spans[1:]))   # a tiling
        assert spans[0][0] == 0 and spans[-1][0] + spans[-1][1] == a.size
        assert all(o1 + s1 < o2 for (o1, s1), (o2, _) in zip(a.free, a.free[1:]))  # coalesced
    for off, n in live:
        a.release(off, n)
    assert a.free == [(0, a.size)]


def test_server_places_requests_in_the_arena():
    """A backend with an arena: each admitted request holds a range until it
    completes; ranges of concurrent requests never overlap; all are returned."""
    class ArenaBackend(FakeBackend):
        def __init__(self, size):
            super().__init__()
            self.arena = serve.Arena(size, small=1 << 16)
            self.held = {}

        def reserve(self, nbytes):
            off = self.arena.alloc(nbytes)
            return None if off is None else (off, nbytes)

        def release(self, token):
            self.arena.release(*token)

        def prepare(self, job, slot, token=None):
            assert token is not None
            lo, hi = token[0], token[0] + self.arena._round(token[1])
            assert all(hi <= o or lo >= e for o, e in self.held.values())
            h = super().prepare(job, slot)
            self.held[h.serial] = (lo, hi)
            return h

        def finish(self, h):
            del self.held[h.serial]
            return super().finish(h)

    reqs = [_disp_request(16 + 40 * (i % 4), 0.0) for i in range(12)]
    size = 3 * runtime.job_bytes(M.api.plan(reqs[3].fn, *reqs[3].args))
    be = ArenaBackend(size)
    out = serve.Server(slots=4, backend=be, budget_bytes=size).run(reqs, keep=False, realtime=False)
    assert all(s.error is None for s in out)
    assert be.arena.free == [(0, be.arena.size)] and not be.held


class FakeQueueBackend(FakeBackend):
    """The queue mode's batches: each completes 1 ms after issue."""

    def __init__(self):
        super().__init__()
        self.batches = []

    def issue_batch(self, items):
        self.batches.append([(h.serial, lo, hi, h.job.n_src * h.job.K) for h, lo, hi in items])
        return _Event(time.perf_counter() + 1e-3)


def test_queue_mode_batches_least_work_first_within_budget():
    """Queue mode: every request's rows are issued once, each batch holds at
    most one chunk per request and stays within batch_pairs (or is a single
    chunk), and within a batch the requests with the least work left come
    first."""
    be = FakeQueueBackend()
    reqs = [_disp_request(n, 0.0) for n in (200, 24, 120, 40, 8)]
    srv = serve.Server(slots=5, chunk_pairs=16 * 20, batch_pairs=16 * 50, backend=be, queue=True,
                       least_work_first=True)
    out = srv.run(reqs, realtime=False)
    assert all(s.error is None for s in out)
    assert srv.batches == len(be.batches) > 1
    rows = {}
    for b in be.batches:
        serials = [h for h, _, _, _ in b]
        assert len(set(serials)) == len(serials)
        pairs = sum((hi - lo) * c for _, lo, hi, c in b)
        assert pairs <= 16 * 50 or len(b) == 1
        for h, lo, hi, _ in b:
            assert hi - lo <= 20
            rows.setdefault(h, []).append((lo, hi))
    assert sorted(sum(hi - lo for lo, hi in v) for v in rows.values()) == sorted((200, 24, 120, 40, 8))
    # the first batch starts with the smallest request (8 rows), then 24
    assert [be.batches[0][i][2] - be.batches[0][i][1] for i in range(2)] == [8, 20]


def test_queue_mode_default_takes_every_request_in_admission_order():
    """The queue mode's default batch: the next chunk of every active
    request, in admission order, with no pair budget."""
    be = FakeQueueBackend()
    reqs = [_disp_request(n, 0.0) for n in (120, 8, 40)]
    srv = serve.Server(slots=3, chunk_pairs=16 * 20, backend=be, queue=True)
    out = srv.run(reqs, realtime=False)
    assert all(s.error is None for s in out)
    assert [h for h, _, _, _ in be.batches[0]] == [1, 2, 3]
    assert [hi - lo for _, lo, hi, _ in be.batches[0]] == [20, 8, 20]
