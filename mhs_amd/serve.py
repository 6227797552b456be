"""Continuous-batching request loop for assembly requests on one GPU.

A request is one public matrix call (``disp_matrix``, ``traction_matrix``,
``interaction_matrix``, ...) with its arguments, an arrival time and a
precision.  The loop:

* **admits** requests first-come first-served while a slot (a HIP stream of
  its own) is free and the request's predicted device bytes (outputs,
  scratch, fold list, inputs: :func:`runtime.job_bytes`) fit in what is left
  of the HBM budget; a request that does not fit yet is passed by later ones
  that do for a bounded time (``bypass_s``), after which the queue waits for
  it; a request larger than the whole budget is refused with the same
  ``MemoryError`` arithmetic as the byte ceiling;
* **interleaves** the admitted requests: each is cut into observer-row chunks
  of bounded pair count and the loop issues one chunk per active request in
  turn (at most ``depth`` in flight each), so a request that arrives while a
  large one is running starts on the device at once and its kernels overlap
  the large one's on another stream -- or, in the queue mode, the next chunk
  of every active request goes into one batch whose passes are each drained
  from a device work queue by one persistent launch (``mhs_assemble_batch``);
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
    def __init__(self, index, prep, chunks, slot, nbytes, token):
        self.index, self.prep, self.slot, self.bytes, self.token = index, prep, slot, nbytes, token
        self.pending = collections.deque(chunks)
        self.inflight = collections.deque()
        self.cols = prep.job.n_src * prep.job.K                            # element columns
        self.left = sum((hi - lo) for lo, hi in chunks) * self.cols        # pairs not yet issued


#: Device memory the tensor cache may never take while serving: the HIP
#: runtime allocates each hardware queue's scratch (the closed-form kernels
#: hold up to 12.6 KB per lane, gigabytes per queue at full occupancy) from
#: what the cache has not reserved, and aborts the queue when that runs out
#: (seen with 308 MB left: results of tens of GB had filled the cache).
SCRATCH_HEADROOM_BYTES = 32 << 30
#: Unused bytes of the tensor cache above which a completion returns them.
TRIM_ABOVE_BYTES = 8 << 30
#: Cached blocks at least this large (MiB) are neither split for smaller
#: tensors nor kept when a larger request cannot be placed.
MAX_SPLIT_MB = 1024


class Arena:
    """Byte ranges of one device buffer, first fit: requests of at least
    ``small`` bytes from the low end, smaller ones from the high end (so
    short-lived small results do not split the space large ones need);
    freed ranges coalesce with their neighbours; ``align``-byte aligned.
    Pure bookkeeping: :class:`DeviceArena` holds the memory."""

    def __init__(self, nbytes: int, small: int = 1 << 30, align: int = 256):
        self.align = int(align)
        self.size = int(nbytes) - int(nbytes) % self.align
        self.small = int(small)
        self.free = [(0, self.size)] if self.size > 0 else []    # (offset, bytes), sorted, disjoint

    def _round(self, n: int) -> int:
        return max(self.align, -(-int(n) // self.align) * self.align)

    def alloc(self, nbytes: int):
        """Offset of a new range of ``nbytes`` (rounded up), or None."""
        n = self._round(nbytes)
        idx = range(len(self.free))
        for k in (reversed(idx) if n < self.small else idx):
            off, size = self.free[k]
            if size < n:
                continue
            if n < self.small:             # from the high end of the block
                self.free[k] = (off, size - n)
                at = off + size - n
            else:
                self.free[k] = (off + n, size - n)
                at = off
            if self.free[k][1] == 0:
                del self.free[k]
            return at
        return None

    def release(self, off: int, nbytes: int) -> None:
        n = self._round(nbytes)
        k = 0
        while k < len(self.free) and self.free[k][0] < off:
            k += 1
        if (k > 0 and self.free[k - 1][0] + self.free[k - 1][1] > off) or \
                (k < len(self.free) and off + n > self.free[k][0]):
            raise ValueError(f"range [{off}, {off + n}) is not allocated")
        self.free.insert(k, (off, n))
        if k + 1 < len(self.free) and off + n == self.free[k + 1][0]:
            self.free[k] = (off, n + self.free[k + 1][1])
            del self.free[k + 1]
        if k > 0 and self.free[k - 1][0] + self.free[k - 1][1] == off:
            self.free[k - 1] = (self.free[k - 1][0], self.free[k - 1][1] + self.free[k][1])
            del self.free[k]

    @property
    def free_bytes(self) -> int:
        return sum(s for _, s in self.free)

    @property
    def largest(self) -> int:
        return max((s for _, s in self.free), default=0)


class DeviceArena(Arena):
    """An :class:`Arena` over one device buffer of the tensor cache: the
    serving loop's results and scratch, allocated once (one hipMalloc
    instead of one per request, and no device-wide synchronisation of a
    free while other requests run)."""

    def __init__(self, torch, dev, nbytes: int):
        super().__init__(nbytes)
        self.buf = torch.empty(self.size, dtype=torch.uint8, device=dev)

    def bump(self, off: int, nbytes: int):
        """An allocator handing out consecutive aligned slices of the range
        [off, off + nbytes) as uint8 tensors (for runtime.Prepared)."""
        end = off + nbytes
        at = [off]

        def take(n: int):
            o = at[0]
            at[0] = o + self._round(n)
            if at[0] > end:
                raise MemoryError(f"arena range of {nbytes} bytes exhausted")
            return self.buf[o:o + n]
        return take


_ARENAS: dict = {}


def device_arena(dev: runtime.Device, nbytes: int) -> DeviceArena:
    """The device's serving arena, created at the first call with ``nbytes``
    bytes and shared by every backend on the device after that."""
    key = dev.dev.index
    if key not in _ARENAS:
        _ARENAS[key] = DeviceArena(dev.torch, dev.dev, nbytes)
    return _ARENAS[key]


def release_arenas() -> None:
    """Drop the serving arenas (their memory returns to the tensor cache)."""
    _ARENAS.clear()


class GpuBackend:
    """Device side of the loop: one stream (default priority) per slot, the
    HIP kernels through :class:`runtime.Prepared`, completion by events.

    The tensor cache is capped at all but ``headroom`` bytes of the device
    (torch's per-process memory fraction: past it the cache first returns
    its unused blocks, then raises), and after requests complete its unused
    blocks are returned whenever they exceed ``trim_above`` bytes.  Both are
    process-wide settings of the caching allocator, as is MAX_SPLIT_MB.
    Requests whose results do not leave the loop take their device buffers
    from the device's serving arena (:func:`device_arena`, the budget's
    size, allocated at the first such request).

    With ``tune`` False (the default) a request uses the device's tuned
    mixed rule and far-pass tiles for its class if an earlier run tuned
    them, else the library defaults, so the loop never waits on a tuner;
    with ``tune`` True the first request of each class tunes them (and
    synchronises its stream), as a warm-up does."""

    def __init__(self, slots: int, dev: runtime.Device | None = None, headroom: int = SCRATCH_HEADROOM_BYTES,
                 tune: bool = False, queue: bool = False):
        self.dev = dev or runtime.device()
        self.queue = bool(queue)
        torch = self.torch = self.dev.torch
        self.total = int(torch.cuda.get_device_properties(self.dev.dev).total_memory)
        self.headroom = int(headroom)
        if not 0 < self.headroom < self.total // 2:
            raise ValueError(f"headroom must be in (0, {self.total // 2}) bytes, got {headroom}")
        # results of tens of GB come and go: keep the tensor cache from
        # carving small tensors out of a finished result's block (which pins
        # the whole segment) and let it return such blocks when a large
        # request does not fit
        torch._C._accelerator_setAllocatorSettings(f"max_split_size_mb:{MAX_SPLIT_MB}")
        torch.cuda.set_per_process_memory_fraction((self.total - self.headroom) / self.total, self.dev.dev)
        self.trim_above = TRIM_ABOVE_BYTES
        # the queue mode runs every request on one stream, in batches
        self.streams = [torch.cuda.Stream(device=self.dev.dev) for _ in range(1 if self.queue else slots)]
        self.tune = bool(tune)
        self.trims = 0          # times the cache was returned to the device

    def default_budget(self) -> int:
        """Bytes of admitted requests at once: 80 % of the device, within
        the cache's cap (the serving arena's size)."""
        return min(int(0.8 * self.total), self.total - self.headroom)

    def reserve(self, nbytes: int):
        """A range of the serving arena for a request's device buffers, or
        None when no free range is large enough yet."""
        arena = device_arena(self.dev, self.default_budget())
        n = int(nbytes) + 4096
        off = arena.alloc(n)
        return None if off is None else (off, n)

    def release(self, token) -> None:
        device_arena(self.dev, self.default_budget()).release(*token)

    def trim(self) -> None:
        """Return the tensor cache's unused blocks to the device when they
        exceed ``trim_above`` bytes (requests of many sizes leave it holding
        blocks that fit no later request)."""
        cuda = self.torch.cuda
        if cuda.memory_reserved(self.dev.dev) - cuda.memory_allocated(self.dev.dev) > self.trim_above:
            cuda.empty_cache()
            self.trims += 1

    def prepare(self, job: runtime.Job, slot: int, token=None):
        alloc = device_arena(self.dev, self.default_budget()).bump(*token) if token is not None else None
        stream = self.streams[0] if self.queue else self.streams[slot]
        prep = runtime.Prepared(job, stream=stream, dev=self.dev, alloc=alloc)
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

    def issue_batch(self, items):
        """Enqueue one batch of (prep, lo, hi) row ranges (distinct jobs) on
        the queue stream: one persistent launch per kernel instance drains
        every range's tiles (runtime.Batch); returns it (query() is True
        once done)."""
        return runtime.Batch(items, self.streams[0], self.dev)

    def finish(self, prep):
        """(out, out2) of a completed request; raises on a fold overflow."""
        prep.check_fold()
        return prep.final, prep.final2


class Server:
    """The request loop.  ``slots`` concurrent requests, each on its own
    stream; ``budget_bytes`` of device memory for admitted work (default 80 %
    of the device); chunks of at most ``chunk_pairs`` pairs; ``depth`` chunks
    in flight per request; a request that does not fit in the budget is
    passed by later ones for at most ``bypass_s`` seconds after its arrival,
    and after that by requests smaller than ``small_bytes`` only (default
    1/64 of the budget; bypass_s = small_bytes = 0: strict first come,
    first served).

    ``queue=True``: the admitted requests share one stream and run in
    batches: each batch takes the next chunk (default 2^22 pairs) of every
    active request in admission order (``least_work_first``: those with the
    least work left first), up to ``batch_pairs`` pairs (default no limit;
    at least one chunk), and enqueues them through
    ``mhs_assemble_batch``, where each tiled pass of every chunk goes into a
    device work queue of (request, observer-row tile, column tile) items
    drained by ONE persistent launch per kernel instance; ``depth`` batches
    are queued on the stream at a time.  ``slots`` then bounds the requests
    admitted at once."""

    def __init__(self, slots: int = 4, budget_bytes: int | None = None, chunk_pairs: int | None = None,
                 depth: int = 2, backend=None, bypass_s: float = 2.0, small_bytes: int | None = None,
                 queue: bool = False, batch_pairs: int | None = None, least_work_first: bool = False):
        self.slots = int(slots)
        self.queue = bool(queue)
        self.least_work_first = bool(least_work_first)
        if chunk_pairs is None:
            chunk_pairs = (1 << 22) if self.queue else (1 << 25)
        self.batch_pairs = int(batch_pairs if batch_pairs is not None else 1 << 62)
        self.bypass_s = float(bypass_s)
        self.backend = backend if backend is not None else GpuBackend(self.slots, queue=self.queue)
        if self.queue and not hasattr(self.backend, "issue_batch"):
            raise ValueError("queue=True needs a backend with issue_batch")
        self.budget = int(budget_bytes if budget_bytes is not None else self.backend.default_budget())
        self.small_bytes = int(self.budget // 64 if small_bytes is None else small_bytes)
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
        released: a consumer that copies or reduces each matrix as it lands
        (it must have finished reading them when it returns).  Unless results
        are kept (``keep`` without ``consume``), a backend with an arena
        (``reserve``) places each request's device buffers in it."""
        be = self.backend
        # device buffers from the backend's arena unless results leave the loop
        arena = hasattr(be, "reserve") and (consume is not None or not keep)
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        served = [Served(i, r.tag, float(r.arrival)) for i, r in enumerate(requests)]
        waiting: list = []
        active: list[_Active] = []
        free_slots = list(range(self.slots))
        in_use = 0
        nxt = 0
        # host seconds per step of the loop (the loop is one host thread:
        # while it plans or prepares, nothing else is admitted or issued)
        st = self.stats = dict.fromkeys(("plan", "prepare", "issue", "finish", "trim", "sleep"), 0.0)
        clock = time.perf_counter
        rounds = collections.deque()        # queue mode: batches in flight, in stream order
        self.batches = 0
        t0 = clock()
        while nxt < len(order) or waiting or active:
            now = time.perf_counter() - t0
            while nxt < len(order) and (not realtime or requests[order[nxt]].arrival <= now):
                i = order[nxt]
                nxt += 1
                tp = clock()
                try:
                    job = self._plan(requests[i])
                    waiting.append((i, job, runtime.job_bytes(job)))
                except Exception as e:            # validation errors are the request's result
                    served[i].error = f"{type(e).__name__}: {e}"
                    served[i].done = served[i].admitted = time.perf_counter() - t0
                st["plan"] += clock() - tp
            # admission: first come, first served, except that a request
            # that does not fit in the budget yet lets later ones that fit
            # pass it (first fit) until it has waited ``bypass_s``; then the
            # queue holds for it while running requests drain, except for
            # small requests (below ``small_bytes``), which still pass
            k = 0
            held = False
            now = time.perf_counter() - t0
            while free_slots and k < len(waiting):
                i, job, nbytes = waiting[k]
                if nbytes > self.budget:
                    del waiting[k]
                    served[i].error = (f"MemoryError: needs {nbytes} device bytes, more than the "
                                       f"{self.budget}-byte serving budget")
                    served[i].done = served[i].admitted = time.perf_counter() - t0
                    continue
                if held and nbytes >= self.small_bytes:
                    k += 1
                    continue
                token = None
                fits = in_use + nbytes <= self.budget
                if fits and arena:
                    token = be.reserve(nbytes)
                    fits = token is not None
                if not fits:
                    if now - served[i].arrival >= self.bypass_s:
                        held = True
                    k += 1
                    continue
                del waiting[k]
                slot = free_slots.pop(0)
                tp = clock()
                prep = be.prepare(job, slot, token) if token is not None else be.prepare(job, slot)
                st["prepare"] += clock() - tp
                served[i].admitted = time.perf_counter() - t0
                served[i].pairs = job.n_obs * job.n_src * job.K
                served[i].bytes = nbytes
                in_use += nbytes
                active.append(_Active(i, prep, prep.chunks(self.chunk_pairs), slot, nbytes, token))
            # issue: one chunk per active request in turn
            issued = False
            tp = clock()
            if self.queue:
                # batches: the next chunk of every active request, all of
                # them drained by one persistent launch per kernel instance;
                # up to ``depth`` batches queued on the stream
                while rounds and rounds[0].query():
                    rounds.popleft()
                while len(rounds) < self.depth:
                    # the active requests in admission order (or with the
                    # least work left first), one chunk each while the
                    # batch has room (at least one chunk)
                    take = [a for a in active if a.pending]
                    if self.least_work_first:
                        take.sort(key=lambda a: a.left)
                    if not take:
                        break
                    items, room = [], self.batch_pairs
                    for a in take:
                        lo, hi = a.pending[0]
                        pairs = (hi - lo) * a.cols
                        if items and pairs > room:
                            continue
                        a.pending.popleft()
                        a.left -= pairs
                        room -= pairs
                        items.append((a, lo, hi))
                    b = be.issue_batch([(a.prep, lo, hi) for a, lo, hi in items])
                    rounds.append(b)
                    for a, _, _ in items:
                        a.inflight.append(b)
                    issued = True
                    self.batches += 1
            else:
                for a in active:
                    while a.pending and len(a.inflight) < self.depth:
                        lo, hi = a.pending.popleft()
                        a.inflight.append(be.issue(a.prep, lo, hi))
                        issued = True
            st["issue"] += clock() - tp
            # completions
            finished = []
            for a in active:
                while a.inflight and a.inflight[0].query():
                    a.inflight.popleft()
                if not a.pending and not a.inflight:
                    finished.append(a)
            tp = clock()
            for a in finished:
                active.remove(a)
                s = served[a.index]
                s.done = time.perf_counter() - t0
                try:
                    res = be.finish(a.prep)
                    if consume is not None:
                        s.out = consume(a.index, *res)
                    elif keep:
                        s.out, s.out2 = res
                    # no reference may outlive the request: its bytes are
                    # released from the budget below
                    del res
                except RuntimeError as e:
                    s.error = str(e)
                free_slots.append(a.slot)
                in_use -= a.bytes
                a.prep = None
                if a.token is not None:
                    be.release(a.token)
            st["finish"] += clock() - tp
            if finished and hasattr(be, "trim"):
                tp = clock()
                be.trim()
                st["trim"] += clock() - tp
            if not issued and not finished:
                # nothing to do until a chunk completes or the next arrival:
                # poll at a fine grain so that neither is noticed late
                wait = 2e-4
                if not active and nxt < len(order) and realtime:
                    wait = max(0.0, min(0.01, requests[order[nxt]].arrival - (time.perf_counter() - t0)))
                if wait > 0:
                    tp = clock()
                    time.sleep(wait)
                    st["sleep"] += clock() - tp
        return served


def summary(served: list[Served]) -> dict:
    """p50 / p99 latency and pair-evals/s over the requests that completed,
    with latency split into the wait for admission (admitted - arrival) and
    the run (done - admitted)."""
    ok = [s for s in served if s.error is None]
    lat = np.array([s.latency for s in ok]) if ok else np.zeros(1)
    wait = np.array([s.admitted - s.arrival for s in ok]) if ok else np.zeros(1)
    run = lat - wait
    first = min(s.arrival for s in served) if served else 0.0
    last = max(s.done for s in ok) if ok else 0.0
    pairs = sum(s.pairs for s in ok)
    return {"requests": len(served), "completed": len(ok), "refused": len(served) - len(ok),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": float(last - first), "pairs": int(pairs),
            "pair_evals_per_s": float(pairs / max(last - first, 1e-9)),
            "wait_p50_s": float(np.percentile(wait, 50)), "wait_p99_s": float(np.percentile(wait, 99)),
            "run_p50_s": float(np.percentile(run, 50)), "run_p99_s": float(np.percentile(run, 99))}


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
