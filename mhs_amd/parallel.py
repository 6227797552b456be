"""Block drivers with the interface of the reference's process-pool helpers
``by_source`` and ``by_rows`` (src/mhs/parallel.py:132-330), so callers
written against them run unchanged.

The reference spreads one assembly over spawned worker processes, because
its per-source loop is many small numpy calls.  Here a single call already
assembles the whole matrix on the GPU, so these drivers run their blocks in
THIS process, one after another, each block one device assembly.  What they
keep of the reference's contract:

* the signatures and keywords (``workers`` and ``blas_threads`` are accepted;
  there is no pool and no BLAS on this path, so they change nothing);
* a small probe block evaluated first to learn the result's shape and bytes
  per source (or per row), then the remaining blocks;
* ``out=``: any ndarray of the exact stitched shape -- a ``numpy.memmap``
  included, which is how a matrix larger than host RAM is filled, block by
  block -- else a new array; a wrong shape raises ``ValueError`` with the
  reference's message;
* stitching along the slip-DOF axis (``source_axis``: -2 for the
  ``(..., n_dof, 3)`` matrices, 1 for ``interaction_matrix``) or along the
  leading receiver-row axis.

Blocks are sized so that one block's result stays under ``block_bytes``
(default ``defaults.DEVICE_BLOCK_BYTES``): the device holds one block's
output at a time, whatever the size of the whole matrix.  Every entry of
these matrices depends on its own (receiver, source) pair only, so the
stitched matrix is bitwise the one a single call returns.

``by_source`` cuts columns, which are disjoint only in the discontinuous
layout; the continuous layout (shared nodes) is cut by receiver rows with
``by_rows``, as in the reference.

:func:`stream_rows` is the out-of-core path of this package's own calls:
receiver-row blocks of one validated call, each block's device-to-host copy
overlapping the next block's assembly, into a host array or memmap.
"""
from __future__ import annotations

import dataclasses
import os
from contextlib import contextmanager

import numpy as np

from . import api, defaults, runtime

#: Threading variables of the BLAS libraries (kept for callers that use the
#: reference's ``pinned_blas``; this package's device path uses no BLAS).
THREAD_ENV = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS",
              "VECLIB_MAXIMUM_THREADS", "NUMEXPR_NUM_THREADS")

#: Rows of ``by_rows``'s probe block (parallel.py:219-224 of the reference).
PROBE_ROWS = 8

#: Fewest rows in one block of ``plan_row_blocks`` (parallel.py:226-239).
ROWS_PER_BLOCK_MIN = 256


@contextmanager
def pinned_blas(n_threads: int = 1):
    """Set every ``THREAD_ENV`` variable to ``n_threads`` for the duration,
    then restore the environment exactly: variables that were absent are
    removed again rather than left set (parallel.py:82-100)."""
    saved = {k: os.environ.get(k) for k in THREAD_ENV}
    try:
        for k in THREAD_ENV:
            os.environ[k] = str(int(n_threads))
        yield
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _ranges(n: int, n_blocks: int) -> list[tuple[int, int]]:
    edges = np.linspace(0, n, n_blocks + 1).round().astype(int)
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


def plan_chunks(n_src: int, workers: int, bytes_per_src: int,
                budget: int | None = None) -> list[tuple[int, int]]:
    """Source ranges as the reference plans them for its pool
    (parallel.py:103-121): at least one per worker, more while the blocks in
    flight (one per worker) would exceed ``budget`` (default
    ``defaults.PARALLEL_INFLIGHT_BYTES``)."""
    n_src, workers = int(n_src), max(1, int(workers))
    if n_src <= 0:
        return []
    cap = int(defaults.PARALLEL_INFLIGHT_BYTES if budget is None else budget)
    per_chunk = max(1, cap // max(workers * max(int(bytes_per_src), 1), 1))
    n_chunks = min(max(workers, -(-n_src // per_chunk)), n_src)
    return _ranges(n_src, n_chunks)


def plan_row_blocks(n_rows: int, workers: int, bytes_per_row: int,
                    budget: int | None = None) -> list[tuple[int, int]]:
    """Row ranges as the reference plans them for its pool
    (parallel.py:242-258): sized against the in-flight budget, never under
    ``ROWS_PER_BLOCK_MIN`` rows, one per worker when there are enough rows."""
    n_rows, workers = int(n_rows), max(1, int(workers))
    if n_rows <= 0:
        return []
    cap = int(defaults.PARALLEL_INFLIGHT_BYTES if budget is None else budget)
    per_block = max(ROWS_PER_BLOCK_MIN, cap // max(workers * max(int(bytes_per_row), 1), 1))
    n_blocks = max(1, -(-n_rows // per_block))
    if n_blocks < workers and n_rows >= workers * ROWS_PER_BLOCK_MIN:
        n_blocks = workers
    return _ranges(n_rows, min(n_blocks, n_rows))


def _device_blocks(n: int, bytes_per_item: int, block_bytes: int | None) -> list[tuple[int, int]]:
    """Consecutive ranges of ``n`` items whose results fit ``block_bytes``."""
    cap = int(defaults.DEVICE_BLOCK_BYTES if block_bytes is None else block_bytes)
    per = max(1, cap // max(int(bytes_per_item), 1))
    return _ranges(n, max(1, -(-n // per)))


def _target(out, shape, dtype):
    if out is None:
        return np.empty(shape, dtype=dtype)
    if out.shape != shape:
        raise ValueError(f"out has shape {out.shape}, expected {shape}")
    return out


def by_source(call, tris, *, workers: int | None = None, out=None,
              source_axis: int = -2, tris_kw: str | None = None,
              blas_threads: int = 1, chunks: int | None = None,
              block_bytes: int | None = None):
    """Evaluate ``call`` on blocks of ``tris`` and stitch the results along
    the slip-DOF axis (the reference's ``by_source``, parallel.py:132-238).

    ``call`` takes one ``tris`` sub-array (as ``tris_kw=`` when given, else
    positionally) and must depend on the sources only through it: an entry
    point that derives its receivers from ``tris``, as ``interaction_matrix``
    does, needs ``obs_tris=`` (or ``obs_pts=``) bound to the full set.
    ``source_axis`` is the DOF axis of the result: -2 for the
    ``(..., n_dof, 3)`` matrices, 1 for ``interaction_matrix``."""
    del workers, blas_threads, chunks        # one device, in this process
    tris = np.ascontiguousarray(np.asarray(tris, float).reshape(-1, 3, 3))
    n_src = tris.shape[0]

    def run(lo, hi):
        sub = tris[lo:hi]
        return np.asarray(call(**{tris_kw: sub}) if tris_kw else call(sub))

    probe_hi = min(n_src, 1)
    first = run(0, probe_hi)
    axis = source_axis if source_axis >= 0 else first.ndim + source_axis
    dof_per_src = first.shape[axis] // max(probe_hi, 1)
    bytes_per_src = first.nbytes // max(probe_hi, 1)
    shape = list(first.shape)
    shape[axis] = dof_per_src * n_src
    buf = _target(out, tuple(shape), first.dtype)

    def write(lo, hi, block):
        sl = [slice(None)] * buf.ndim
        sl[axis] = slice(lo * dof_per_src, hi * dof_per_src)
        buf[tuple(sl)] = block

    write(0, probe_hi, first)
    del first
    for a, b in _device_blocks(n_src - probe_hi, bytes_per_src, block_bytes):
        write(probe_hi + a, probe_hi + b, run(probe_hi + a, probe_hi + b))
    return buf


def by_rows(call, n_rows: int, *, workers: int | None = None, out=None,
            rows_kw: str = "rows", blas_threads: int = 1,
            block_bytes: int | None = None):
    """Evaluate ``call(rows=(lo, hi))`` on receiver-row blocks and stitch
    them along the leading axis (the reference's ``by_rows``,
    parallel.py:267-330): the cut for the continuous layout, whose columns
    are shared between elements but whose rows are disjoint."""
    del workers, blas_threads                # one device, in this process
    n_rows = int(n_rows)
    if n_rows <= 0:
        raise ValueError(f"n_rows must be positive, got {n_rows}")

    def run(lo, hi):
        return np.asarray(call(**{rows_kw: (lo, hi)}))

    probe_hi = min(n_rows, PROBE_ROWS)
    first = run(0, probe_hi)
    bytes_per_row = first.nbytes // max(first.shape[0], 1)
    buf = _target(out, (n_rows,) + first.shape[1:], first.dtype)
    buf[:probe_hi] = first
    del first
    for a, b in _device_blocks(n_rows - probe_hi, bytes_per_row, block_bytes):
        buf[probe_hi + a:probe_hi + b] = run(probe_hi + a, probe_hi + b)
    return buf


def stream_rows(fn, *args, out=None, block_bytes: int | None = None, **kwargs):
    """``fn(*args, **kwargs)`` -- any public matrix call of this package --
    assembled by receiver-row blocks straight into ``out`` (a host ndarray
    or ``numpy.memmap`` of the call's shape, float64 or float32; else a new
    float64 array), with each block's device-to-host copy overlapping the
    next block's assembly.

    The row blocks are cut from the validated call (receivers are
    independent in every layout), each block's result kept under
    ``block_bytes`` of device memory (default
    ``defaults.DEVICE_BLOCK_BYTES``), two pinned host staging buffers, the
    assembly on one stream and the copies on another.  The matrix equals the
    direct call's bitwise.  This is the out-of-core path: neither the device
    nor host RAM ever holds more than two blocks (the output itself may be a
    memmap).  ``parts=True`` returns two matrices and is not streamed."""
    job = api.plan(fn, *args, **kwargs)
    if job.parts:
        raise ValueError("stream_rows streams one matrix; parts=True wants by_rows")
    shape = job.out_shape()
    if out is None:
        buf = np.empty(shape, dtype=np.float64)
    else:
        if not isinstance(out, np.ndarray):
            raise TypeError(f"out must be a numpy.ndarray, got {type(out).__name__}")
        if out.shape != shape:
            raise ValueError(f"out has shape {out.shape}, expected {shape}")
        if out.dtype not in (np.float64, np.float32):
            raise ValueError(f"out must be float64 or float32, got {out.dtype}")
        buf = out
    return stream_job(job, buf, block_bytes)


def stream_job(job, buf, block_bytes: int | None = None):
    """Assemble the validated ``job`` by receiver-row blocks into the host
    array ``buf`` (its output shape, float64 or float32): the engine of
    :func:`stream_rows`, also used by the public calls for an ``out=``
    larger than one device block."""
    shape = job.out_shape()
    n = job.n_obs
    row_elems = int(np.prod(shape[1:]))
    blocks = _device_blocks(n, 8 * row_elems, block_bytes)
    if not blocks:
        return buf
    dev = runtime.device()
    torch = dev.torch
    compute, copy = torch.cuda.Stream(device=dev.dev), torch.cuda.Stream(device=dev.dev)
    most = max(hi - lo for lo, hi in blocks) * row_elems
    staging = [torch.empty(most, dtype=torch.float64, pin_memory=True) for _ in range(2)]

    def rows_of(lo, hi):
        cut = {"obs": job.obs[lo:hi]}
        for name in ("nrm", "rb"):
            if getattr(job, name) is not None:
                cut[name] = getattr(job, name)[lo:hi]
        return dataclasses.replace(job, **cut)

    def land(item):
        lo, hi, host, done, prep = item
        done.synchronize()
        prep.check_fold()
        buf[lo:hi] = host[: (hi - lo) * row_elems].numpy().reshape((hi - lo,) + shape[1:])

    pending = None
    for i, (lo, hi) in enumerate(blocks):
        prep = runtime.Prepared(rows_of(lo, hi), stream=compute, dev=dev)
        with torch.cuda.stream(compute):
            prep.launch()
            assembled = torch.cuda.Event()
            assembled.record(compute)
        copy.wait_event(assembled)
        host = staging[i % 2]
        with torch.cuda.stream(copy):
            host[: (hi - lo) * row_elems].copy_(prep.final.reshape(-1), non_blocking=True)
            done = torch.cuda.Event()
            done.record(copy)
        if pending is not None:
            land(pending)           # the previous block, while this one assembles
        pending = (lo, hi, host, done, prep)
    land(pending)
    return buf
