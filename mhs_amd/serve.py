"""Continuous-batching request loop for assembly requests on one GPU.

A request is one public matrix call (``disp_matrix``, ``traction_matrix``,
``interaction_matrix``, ...) with its arguments, an arrival time and a
precision.  The loop:

* **admits** requests first-come first-served while a slot (a HIP stream of
  its own) is free and the request's predicted device bytes (outputs,
  scratch, fold list, inputs: :func:`runtime.job_bytes`) fit in what is left
  of the HBM budget; a request larger than the whole budget is refused with
  the same ``MemoryError`` arithmetic as the byte ceiling;
* **interleaves** the admitted requests: each is cut into observer-row chunks
  of bounded pair count and the loop issues one chunk per active request in
  turn (at most ``depth`` in flight each), so a request that arrives while a
  large one is running starts on the device at once and its kernels overlap
  the large one's on another stream;
* **keeps results device-resident** (or drops them, for throughput runs) and
  frees the slot and the bytes as soon as a request's last chunk completes,
  admitting the next waiting request.

Chunks only split observer rows, and every row's entries come from the same
kernels whatever else is running, so a request's result is the same alone
and in the mix.  Latency is completion time minus the scheduled arrival time.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from . import api, runtime


@dataclass
class Request:
    """One assembly request: ``fn(*args, **kwargs)`` at ``arrival`` seconds."""
    fn: Callable
    args: tuple = ()
    kwargs: dict = field(default_factory=dict)
    arrival: float = 0.0
    precision: str = "fp64"
    tag: str = ""


@dataclass
class Served:
    """Outcome of one request (times in seconds from the stream start)."""
    index: int
    tag: str
    arrival: float
    admitted: float = float("nan")
    done: float = float("nan")
    pairs: int = 0
    bytes: int = 0
    out: Any = None          # device tensor in the public output layout (keep=True)
    out2: Any = None
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Active:
    def __init__(self, index, prep, chunks, slot, nbytes):
        self.index, self.prep, self.slot, self.bytes = index, prep, slot, nbytes
        self.pending = collections.deque(chunks)
        self.inflight = collections.deque()


#: Device memory kept free of the tensor cache while serving: the HIP runtime
#: allocates each hardware queue's scratch (the closed-form kernels hold up
#: to 12.6 KB per lane, about 1.6 GB per queue at full occupancy) from what
#: the cache has not reserved, and fails the queue when that runs out.
SCRATCH_HEADROOM_BYTES = 16 << 30


class GpuBackend:
    """Device side of the loop: one stream (default priority) per slot, the
    HIP kernels through :class:`runtime.Prepared`, completion by events.

    With ``tune`` False (the default) a request uses the device's tuned
    mixed rule and far-pass tiles for its class if an earlier run tuned
    them, else the library defaults, so the loop never waits on a tuner;
    with ``tune`` True the first request of each class tunes them (and
    synchronises its stream), as a warm-up does."""

    def __init__(self, slots: int, dev: runtime.Device | None = None, headroom: int = SCRATCH_HEADROOM_BYTES,
                 tune: bool = False):
        self.dev = dev or runtime.device()
        self.torch = self.dev.torch
        self.streams = [self.torch.cuda.Stream(device=self.dev.dev) for _ in range(slots)]
        self.total = int(self.torch.cuda.get_device_properties(self.dev.dev).total_memory)
        self.headroom = int(headroom)
        self.tune = bool(tune)
        self.trims = 0          # times the cache was returned to the device

    def default_budget(self) -> int:
        return int(0.8 * self.total)

    def trim(self) -> None:
        """Return the tensor cache's unused blocks to the device when it has
        reserved more than all but ``headroom`` bytes of it (requests of
        many sizes fragment the cache until it holds nearly all of HBM)."""
        if self.torch.cuda.memory_reserved(self.dev.dev) > self.total - self.headroom:
            self.torch.cuda.empty_cache()
            self.trims += 1

    def prepare(self, job: runtime.Job, slot: int):
        prep = runtime.Prepared(job, stream=self.streams[slot], dev=self.dev)
        if not self.tune:
            prep.fix_rule()
            prep.fix_tiles()
        return prep

    def issue(self, prep, lo: int, hi: int):
        """Enqueue rows [lo, hi); returns an object whose query() is True once done."""
        torch = self.torch
        with torch.cuda.stream(prep.stream):
            prep.launch_rows(lo, hi)
            ev = torch.cuda.Event()
            ev.record(prep.stream)
        return ev

    def finish(self, prep):
        """(out, out2) of a completed request; raises on a fold overflow."""
        prep.check_fold()
        return prep.final, prep.final2


class Server:
    """The request loop.  ``slots`` concurrent requests, each on its own
    stream; ``budget_bytes`` of device memory for admitted work (default 80 %
    of the device); chunks of at most ``chunk_pairs`` pairs; ``depth`` chunks
    in flight per request."""

    def __init__(self, slots: int = 4, budget_bytes: int | None = None, chunk_pairs: int = 1 << 25,
                 depth: int = 2, backend=None):
        self.slots = int(slots)
        self.backend = backend if backend is not None else GpuBackend(self.slots)
        self.budget = int(budget_bytes if budget_bytes is not None else self.backend.default_budget())
        self.chunk_pairs = int(chunk_pairs)
        self.depth = int(depth)

    def _plan(self, req: Request) -> runtime.Job:
        # results stay on the device: admitted against the serving budget
        # below, not the host's byte ceiling
        with runtime.precision(req.precision):
            return api.plan_device(req.fn, *req.args, **req.kwargs)

    def run(self, requests: list[Request], keep: bool = True, realtime: bool = True,
            consume: Callable | None = None) -> list[Served]:
        """Serve ``requests`` (arrival times relative to the call); returns one
        :class:`Served` per request, in input order.  ``consume(index, out,
        out2)``, if given, is called as each request completes (after its
        ``done`` time is taken) with its device results, and what it returns
        is kept in ``Served.out`` in place of the results, which are then
        released: a consumer that copies or reduces each matrix as it lands."""
        be = self.backend
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        served = [Served(i, r.tag, float(r.arrival)) for i, r in enumerate(requests)]
        waiting: collections.deque = collections.deque()
        active: list[_Active] = []
        free_slots = list(range(self.slots))
        in_use = 0
        nxt = 0
        t0 = time.perf_counter()
        while nxt < len(order) or waiting or active:
            now = time.perf_counter() - t0
            while nxt < len(order) and (not realtime or requests[order[nxt]].arrival <= now):
                i = order[nxt]
                nxt += 1
                try:
                    job = self._plan(requests[i])
                    waiting.append((i, job, runtime.job_bytes(job)))
                except Exception as e:            # validation errors are the request's result
                    served[i].error = f"{type(e).__name__}: {e}"
                    served[i].done = served[i].admitted = time.perf_counter() - t0
            # admission: first come, first served
            while waiting and free_slots:
                i, job, nbytes = waiting[0]
                if nbytes > self.budget:
                    waiting.popleft()
                    served[i].error = (f"MemoryError: needs {nbytes} device bytes, more than the "
                                       f"{self.budget}-byte serving budget")
                    served[i].done = served[i].admitted = time.perf_counter() - t0
                    continue
                if in_use + nbytes > self.budget:
                    break
                waiting.popleft()
                slot = free_slots.pop(0)
                prep = be.prepare(job, slot)
                served[i].admitted = time.perf_counter() - t0
                served[i].pairs = job.n_obs * job.n_src * job.K
                served[i].bytes = nbytes
                in_use += nbytes
                active.append(_Active(i, prep, prep.chunks(self.chunk_pairs), slot, nbytes))
            # issue: one chunk per active request in turn
            issued = False
            for a in active:
                while a.pending and len(a.inflight) < self.depth:
                    lo, hi = a.pending.popleft()
                    a.inflight.append(be.issue(a.prep, lo, hi))
                    issued = True
            # completions
            finished = []
            for a in active:
                while a.inflight and a.inflight[0].query():
                    a.inflight.popleft()
                if not a.pending and not a.inflight:
                    finished.append(a)
            for a in finished:
                active.remove(a)
                s = served[a.index]
                s.done = time.perf_counter() - t0
                try:
                    out, out2 = be.finish(a.prep)
                    if consume is not None:
                        s.out = consume(a.index, out, out2)
                    elif keep:
                        s.out, s.out2 = out, out2
                except RuntimeError as e:
                    s.error = str(e)
                free_slots.append(a.slot)
                in_use -= a.bytes
                a.prep = None
            if finished and hasattr(be, "trim"):
                be.trim()
            if not issued and not finished:
                # nothing to do until a chunk completes or the next arrival:
                # poll at a fine grain so that neither is noticed late
                wait = 2e-4
                if not active and nxt < len(order) and realtime:
                    wait = max(0.0, min(0.01, requests[order[nxt]].arrival - (time.perf_counter() - t0)))
                if wait > 0:
                    time.sleep(wait)
        return served


def summary(served: list[Served]) -> dict:
    """p50 / p99 latency and pair-evals/s over the requests that completed."""
    ok = [s for s in served if s.error is None]
    lat = np.array([s.latency for s in ok]) if ok else np.zeros(1)
    first = min(s.arrival for s in served) if served else 0.0
    last = max(s.done for s in ok) if ok else 0.0
    pairs = sum(s.pairs for s in ok)
    return {"requests": len(served), "completed": len(ok), "refused": len(served) - len(ok),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": float(last - first), "pairs": int(pairs),
            "pair_evals_per_s": float(pairs / max(last - first, 1e-9))}


# ------------------------------------------------------------ request mix --
def c5_requests(n_req: int = 512, rate: float = 1.0, precision: str = "fp64",
                seed: int | None = None) -> list[Request]:
    """The served stream: ``n_req`` requests with Poisson arrivals at ``rate``
    per second, meshes from :func:`workloads.c5_request` (random-perturbed
    planar faults, n log-uniform in [100, 8000], dip U[20, 90], top depth
    U[0, 5] km, order U{0, 1, 2}); 40 % ``disp_matrix`` on a 32 x 32 surface
    grid over [-30, 30]^2, 30 % ``traction_matrix`` at 2000 observers uniform
    in the fault's box widened by 10 km (z <= 0) with random unit normals,
    30 % ``interaction_matrix`` with the reference's default components
    (receiver strike / dip / normal, source strike / dip / tensile).  Meshes
    are not cut: the order-2 tail's interaction matrices reach 48,000^2 x 9
    entries (166 GB), admitted against the device's memory, not the host's
    byte ceiling."""
    from . import api, workloads
    from .materials import Material
    rng = np.random.default_rng(workloads.SEED + 5 if seed is None else seed)
    mat = Material(30.0, 30.0)
    g = np.linspace(-30.0, 30.0, 32)
    X, Y = np.meshgrid(g, g, indexing="ij")
    grid = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])
    t = 0.0
    out = []
    for i in range(n_req):
        r = workloads.c5_request(rng)
        tris, p, eps = r["tris"], r["order"], r["eps"]
        K = (p + 1) * (p + 2) // 2
        t += rng.exponential(1.0 / rate)
        if r["kind"] == "disp":
            req = Request(api.disp_matrix, (grid, tris, mat, eps), {"order": p})
        elif r["kind"] == "traction":
            lo = tris.reshape(-1, 3).min(axis=0) - 10.0
            hi = tris.reshape(-1, 3).max(axis=0) + 10.0
            obs = np.column_stack([rng.uniform(lo[0], hi[0], 2000), rng.uniform(lo[1], hi[1], 2000),
                                   -rng.uniform(0.0, max(-lo[2], 1.0), 2000)])
            nrm = rng.normal(size=(2000, 3))
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            req = Request(api.traction_matrix, (obs, nrm, tris, mat, eps), {"order": p})
        else:
            req = Request(api.interaction_matrix, (tris, mat, eps), {"order": p})
        req.arrival, req.precision = t, precision
        req.tag = f"{r['kind']} n={tris.shape[0]} p={p}"
        out.append(req)
    return out
