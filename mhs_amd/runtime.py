"""Device execution of one assembly job: buffers, launches, gather, readout.

PyTorch-ROCm supplies device memory and the stream; every arithmetic step on
the pair path is a HIP kernel behind the C-ABI (``_lib``).  There is no CPU
fallback: without a GPU or without the built library this module raises.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _lib, geometry

ROW_SCRATCH_BYTES = 4 << 30      # continuous-layout element-major scratch per launch
FOLD_ITEM_BYTES = 20200          # upper bound of mhs_fold_item_bytes (checked by the ABI tests)
FAR_TILE_CANDIDATES = (8, 16, 32)      # column tiles per block of the far pass the tuner tries
FAR_TILE_CANDIDATES_MIXED = (1, 2, 4, 8)   # the same for the mixed far pass (tiles of 256 sources)
AUTOTUNE_TIE = 0.05                    # candidates this close to the fastest count as tied: the most tiles wins
AUTOTUNE_PAIRS = 1 << 24               # pairs of the sample each candidate is timed on
AUTOTUNE_REPEATS = 2                   # timed launches per candidate (the fastest counts)
autotune = True                         # off: the library default is used


def tile_candidates(mixed: bool):
    return FAR_TILE_CANDIDATES_MIXED if mixed else FAR_TILE_CANDIDATES


#: The mixed mode's far-field rule, as (near_ratio, ((band_hi, order), ...))
#: or (near_ratio, bands, fold_near_ratio), in units of the source's size L
#: (Gauss points per direction: order + the shape order): None lets the
#: autotuner choose among RULE_CANDIDATES for an
#: on-fault interaction and takes the library's default, RULE_SAFE, for
#: everything else.  See include/mhs_amd.h (mhs_rule).
mixed_rule = None

#: The library's default mixed rule: every pair's quadrature error below
#: 1e-7 of the pair's own largest entry, for every form and shape order
#: (worst case measured: the traction of a constant-slip source, 2 x 1
#: triangles at 20-90 deg dip, against 24 x 24 Gauss: at most 5e-10 at 4-6 L
#: with order 5, 7e-9 at 6-16 L with 4, 8e-9 at 16-100 L with 3 and 6e-8
#: beyond with 2).
#: So each row's error is bounded relative to the row, wherever its observer.
RULE_SAFE = (4.0, ((6.0, 5), (16.0, 4), (100.0, 3), (1e300, 2)))

#: Candidate rules of the tuner for the on-fault interaction: a nearer
#: near/far switch moves pairs and fold points from the closed forms to
#: Gauss points, at higher orders close in; a third entry sets the fold's
#: surface points' switch alone (the first band's order beyond it).  Their low orders far away (one
#: point per direction beyond 60 L in the first) are far less accurate per
#: pair than RULE_SAFE,
#: but the receiver rows of an on-fault interaction each hold their own
#: element's near field, which the far pairs' errors (decaying with the
#: stress kernel's 1 / r^3) stay far below: the tuner measures each row's
#: error relative to the row and keeps only candidates within the bound.
RULE_CANDIDATES = (
    (4.0, ((8.0, 3), (60.0, 2), (1e300, 1))),
    (4.0, ((10.0, 3), (100.0, 2), (1e300, 1))),
    (3.0, ((5.0, 4), (10.0, 3), (100.0, 2), (1e300, 1))),
    (3.0, ((5.0, 4), (10.0, 3), (100.0, 2), (1e300, 1)), 2.0),
    (3.0, ((5.0, 4), (10.0, 3), (100.0, 2), (1e300, 1)), 1.5),
    (2.5, ((5.0, 4), (10.0, 3), (100.0, 2), (1e300, 1)), 1.5),
    (2.0, ((5.0, 4), (10.0, 3), (100.0, 2), (1e300, 1)), 1.5),
    (3.0, ((4.0, 5), (8.0, 4), (20.0, 3), (100.0, 2), (1e300, 1))),
    (2.0, ((3.0, 6), (5.0, 5), (10.0, 4), (30.0, 3), (100.0, 2), (1e300, 1))),
    (5.0, ((12.0, 4), (30.0, 3), (150.0, 2), (1e300, 1))),
    RULE_SAFE,
)
#: The tuner's bound on a candidate's error: the largest over the sample rows
#: of max |mixed - FP64| / max |FP64| in the row (the mixed mode's stated
#: bound is 1e-6 per row).
RULE_ERROR_BOUND = 1e-7
#: Candidates within the bound whose sample time is within this fraction of
#: the fastest count as tied, and the most accurate of them wins: the sample
#: separates near-equal rules by less than a full step's run-to-run spread
#: (a rule with a 3 L first band sampled 0.4% faster than the 5 L one with
#: 5x its error, 9.6e-8 against 1.8e-8, and ran the step no faster).
RULE_TIE = 0.01


def pick_rule(table, bound: float = RULE_ERROR_BOUND, tie: float = RULE_TIE, fallback=None):
    """The tuner's choice from ``table`` = {candidate: (error, ms)}: of the
    candidates with error <= bound, those within ``tie`` of the fastest
    time, the one with the least error (ties in error: the faster, then the
    earlier); ``fallback`` when none is within the bound."""
    ok = [(c, e, t) for c, (e, t) in table.items() if e <= bound]
    if not ok:
        return fallback
    fastest = min(t for _, _, t in ok)
    tied = [(e, t, i, c) for i, (c, e, t) in enumerate(ok) if t <= fastest * (1.0 + tie)]
    return min(tied, key=lambda x: x[:3])[3]


def _rule_name(rule) -> str:
    near, bands = rule[0], rule[1]
    fold = f"; fold points near {rule[2]:g} L" if len(rule) > 2 else ""
    return f"near {near:g} L; " + ", ".join(f"n={n} to {hi:g} L" if hi < 1e299 else f"n={n} beyond"
                                           for hi, n in bands) + fold


def rule_struct(rule):
    """A :class:`_lib.Rule` from (near_ratio, [(band_hi, order), ...]) or
    (near_ratio, bands, fold_near_ratio)."""
    r = _lib.Rule()
    if rule is None:
        return r
    near, bands = rule[0], rule[1]
    r.fold_near_ratio = float(rule[2]) if len(rule) > 2 else 0.0
    if not 1 <= len(bands) <= _lib.RULE_BANDS:
        raise ValueError(f"a rule has 1..{_lib.RULE_BANDS} bands")
    bands = list(bands) + [(1e300, bands[-1][1])] * (_lib.RULE_BANDS - len(bands))
    r.near_ratio = float(near)
    for b, (hi, n) in enumerate(bands):
        r.band_hi[b] = float(hi)
        r.order[b] = int(n)
    return r


_ALL_STAGES = _lib.STAGE_PAIR | _lib.STAGE_Q | _lib.STAGE_FOLD

PRECISIONS = ("fp64", "mixed")
_precision = "fp64"


def set_precision(mode: str) -> None:
    """Arithmetic of later calls: "fp64" (default; every branch in double) or
    "mixed" (far-field quadrature points in FP32 with FP64 accumulation; the
    near field, fold and eigenstress stay FP64)."""
    global _precision
    if mode not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}, got {mode!r}")
    _precision = mode


def get_precision() -> str:
    return _precision


class precision:
    """``with mhs_amd.precision("mixed"): ...`` -- scoped :func:`set_precision`."""

    def __init__(self, mode: str):
        if mode not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}, got {mode!r}")
        self.mode, self.prev = mode, None

    def __enter__(self):
        self.prev = get_precision()
        set_precision(self.mode)
        return self

    def __exit__(self, *exc):
        set_precision(self.prev)
        return False


@dataclass
class Job:
    """A validated assembly request (everything the kernels need, host side)."""
    kind: str                   # disp | stress | eigen | traction | interaction
    order: int
    obs: np.ndarray             # (n_obs, 3)
    tris: np.ndarray            # (n_src, 3, 3)
    eps: np.ndarray             # (n_src,)
    mu: float
    lam: float
    fold: bool = True
    subtract: bool = True
    parts: bool = False
    strain: bool = False
    mixed: bool = False         # FP32 far-field points, FP64 accumulation
    nrm: np.ndarray | None = None      # (n_obs, 3)
    rb: np.ndarray | None = None       # (n_obs, n_a, 3)
    sb: np.ndarray | None = None       # (n_src | n_nodes, n_b, 3)
    ids: np.ndarray | None = None      # (n_src, K) continuous columns
    n_nodes: int = 0
    # receivers are the source mesh's own collocation points (or shared
    # nodes): every row holds its own element's near field
    self_rows: bool = False
    extra: dict = field(default_factory=dict)

    @property
    def K(self) -> int:
        return (self.order + 1) * (self.order + 2) // 2

    @property
    def split(self) -> bool:
        """The stress form keeps the eigenstress in a second buffer: for
        parts=True, and for the elastic stress, formed at the end as the
        finished total minus the finished eigenstress (matrices.py:287-294)."""
        return self.kind == "stress" and (self.parts or self.subtract)

    @property
    def n_obs(self) -> int:
        return int(self.obs.shape[0])

    @property
    def n_src(self) -> int:
        return int(self.tris.shape[0])

    @property
    def n_dof(self) -> int:
        return self.n_nodes if self.ids is not None else self.n_src * self.K

    @property
    def n_a(self) -> int:
        return int(self.rb.shape[1]) if self.rb is not None else 0

    @property
    def n_b(self) -> int:
        return int(self.sb.shape[1]) if self.sb is not None else 0

    def out_shape(self, n_cols=None):
        n = self.n_dof if n_cols is None else n_cols
        if self.kind in ("disp", "traction"):
            return (self.n_obs, 3, n, 3)
        if self.kind in ("stress", "eigen"):
            return (self.n_obs, 3, 3, n, 3)
        return (self.n_obs, n, self.n_a, self.n_b)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("mhs_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


class Device:
    """Per-process device state: the library, Gauss tables, material coefficients."""

    def __init__(self, index: int = 0):
        torch = _torch()
        self.torch = torch
        self.dev = torch.device("cuda", index)
        self.lib = _lib.load()
        gx, gw = geometry.gauss_tables()
        self.gl_x = torch.from_numpy(gx).to(self.dev)
        self.gl_w = torch.from_numpy(gw).to(self.dev)
        self.gq = torch.from_numpy(geometry.gauss_triangle_tables()).to(self.dev)
        self._coef = {}
        self.far_tiles = {}     # autotuner cache: (order, outputs, mixed, log2 columns) -> tiles
        self.tune_times = {}    # the same key -> {tiles: ms of the sample's far pass}
        self.rules = {}         # mixed far-field rule chosen per tune key
        self.rule_table = {}    # the same key -> {rule: {relmax, ms}} of every candidate

    def coef(self, mu: float, lam: float):
        key = (float(mu), float(lam))
        t = self._coef.get(key)
        if t is None:
            host = np.zeros(self.lib.mhs_ncoef())
            _lib.check(self.lib.mhs_fill_material(key[0], key[1], host.ctypes.data), self.lib)
            t = self.torch.from_numpy(host).to(self.dev)
            self._coef[key] = t
        return t

    def put(self, a, dtype=None):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype or np.float64)
        return self.torch.from_numpy(a).to(self.dev)

    def upload(self, arrays, stream):
        """Host arrays to the device in one transfer on ``stream``: packed at
        256-byte offsets into one pinned buffer (from torch's pinned cache,
        which keeps it until the copy has completed), one asynchronous copy
        into one device buffer; returns device views in order (None for
        None).  The host does not wait for the copy."""
        torch = self.torch
        items, off = [], 0
        for a in arrays:
            if a is None:
                items.append(None)
                continue
            a = np.ascontiguousarray(a)
            items.append((off, a))
            off += -(-a.nbytes // 256) * 256
        total = max(off, 256)
        pinned = torch.empty(total, dtype=torch.uint8, pin_memory=True)
        buf = pinned.numpy()
        for it in items:
            if it is not None:
                o, a = it
                buf[o:o + a.nbytes] = a.reshape(-1).view(np.uint8)
        with torch.cuda.stream(stream):
            dbuf = torch.empty(total, dtype=torch.uint8, device=self.dev)
            dbuf.copy_(pinned, non_blocking=True)
        out = []
        for it in items:
            if it is None:
                out.append(None)
                continue
            o, a = it
            dt = torch.from_numpy(np.empty(0, a.dtype)).dtype
            out.append(dbuf[o:o + a.nbytes].view(dt).view(a.shape))
        return out


_DEVICE = None


def device() -> Device:
    global _DEVICE
    if _DEVICE is None:
        _DEVICE = Device(0)
    return _DEVICE


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def use_contracted(kind: str, has_sb: bool, n_a: int, n_b: int, order: int, mixed: bool) -> bool:
    """The library's choice between the contracted passes and the tensor
    passes for an interaction launch (csrc/kernels.hip use_contracted): one
    or two components always; more in the mixed mode while the output groups
    (a receiver row with up to two source rows) are no more than the node
    columns per source."""
    if kind != "interaction" or not has_sb:
        return False
    if n_a * n_b <= 2:
        return True
    K = (order + 1) * (order + 2) // 2
    return bool(mixed) and n_a * ((n_b + 1) // 2) <= K


def near_tile_bounds(recs: np.ndarray, K: int) -> np.ndarray:
    """Per 256 element columns (n_cols = n_src K): the bounding box of the
    covered sources' centroids and mirrored centroids (lo xyz, hi xyz), their
    largest L and whether any straddles (the fold), as (tiles, 8) float64:
    the near pass skips a step of columns whose box lies beyond the switch."""
    n_src = recs.shape[0]
    n_t = (n_src * K + 255) // 256
    first = (256 * np.arange(n_t)) // K
    last = np.minimum(n_src, (256 * np.arange(n_t) + 255) // K + 1)
    c = recs[:, geometry.S_C:geometry.S_C + 3]
    ic = recs[:, geometry.S_IC:geometry.S_IC + 3]
    lo_s, hi_s = np.minimum(c, ic), np.maximum(c, ic)
    L = np.maximum(recs[:, geometry.S_L], recs[:, geometry.S_IL])
    st = (recs[:, geometry.S_STRADDLE] != 0.0).astype(float)
    out = np.empty((n_t, 8))
    for j in range(n_t):
        sl = slice(first[j], last[j])
        out[j, 0:3] = lo_s[sl].min(axis=0)
        out[j, 3:6] = hi_s[sl].max(axis=0)
        out[j, 6] = L[sl].max()
        out[j, 7] = st[sl].max()
    return out


def _csr(ids: np.ndarray, n_nodes: int):
    """Element columns of every shared node, ascending: (ptr, idx)."""
    flat = ids.reshape(-1)
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=n_nodes)
    ptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    return ptr, order.astype(np.int64)


class Prepared:
    """A job with its inputs resident in HBM and its outputs allocated: the
    host-side planning happens here, once; :meth:`launch` only enqueues
    kernels (what a benchmark times and what the serving engine replays).

    The inputs go to the device in one asynchronous copy (Device.upload).
    ``alloc(nbytes)``, if given, supplies the large buffers (outputs, the
    continuous layout's element scratch, the fold list and work) as uint8
    device tensors (the serving loop carves them out of its arena); by
    default they come from the tensor cache."""

    def __init__(self, job: Job, stream=None, dev: Device | None = None, alloc=None):
        dev = dev or device()
        self.dev, self.job = dev, job
        torch = dev.torch
        self.stream = stream if stream is not None else torch.cuda.current_stream(dev.dev)
        self.handle = C.c_void_p(self.stream.cuda_stream)

        def empty(shape, dtype=torch.float64):
            if alloc is None:
                return torch.empty(shape, dtype=dtype, device=dev.dev)
            n = int(np.prod(shape)) * torch.empty(0, dtype=dtype).element_size()
            return alloc(n).view(dtype).view(shape)

        K = job.K
        self.n_cols = n_cols = job.n_src * K
        recs = job_records(job)
        strad = (np.flatnonzero(recs[:, geometry.S_STRADDLE] != 0.0).astype(np.int64)
                 if job.fold else np.zeros(0, np.int64))
        self.n_strad = len(strad)
        self.cont = cont = job.ids is not None
        self.kind = _lib.KIND[job.kind]
        self.flags = ((_lib.F_FOLD if job.fold else 0) | (_lib.F_SUBTRACT if job.subtract else 0)
                      | (_lib.F_PARTS if job.parts else 0) | (_lib.F_MIXED if job.mixed else 0))
        self.n_b_launch = 3 if (cont and job.kind == "interaction") else job.n_b
        # the continuous layout's element columns take Cartesian slip
        # (resolved on the node frames by the gather): unit source rows
        sb = (np.broadcast_to(np.eye(3), (job.n_src, 3, 3)) if cont and job.kind == "interaction"
              else None if cont else job.sb)
        # the interaction runs the contracted passes (pointwise far field,
        # closed-form near field, in output groups of one or two components),
        # whose far pass has a tunable launch shape: column tiles per block
        self.contracted = use_contracted(job.kind, sb is not None, job.n_a, self.n_b_launch,
                                         job.order, job.mixed)
        f64 = lambda a: None if a is None else np.asarray(a, np.float64)
        host = [f64(job.obs), recs, f64(job.nrm), f64(job.rb), f64(sb),
                strad if strad.size else None,
                near_tile_bounds(recs, K) if self.contracted else None,
                np.zeros(2, np.int32), np.zeros(1, np.int32)]
        if cont:
            ptr, idx = _csr(job.ids, job.n_nodes)
            host += [ptr, idx, f64(job.sb) if job.kind == "interaction" else None]
        dev_in = dev.upload(host, self.stream)
        (self.obs_t, self.src_t, self.nrm_t, self.rb_t, self.sb_t, self.strad_t, self.near_tiles_t,
         self.fold_cnt, self.fold_err) = dev_in[:9]
        self.coef_t = dev.coef(job.mu, job.lam)
        with torch.cuda.stream(self.stream):
            self.final = empty(job.out_shape())
            self.final2 = empty(job.out_shape()) if job.split else None
            if cont:
                elem_shape = list(job.out_shape(n_cols))
                if job.kind == "interaction":
                    elem_shape[-1] = 3
                self.elem_shape = elem_shape
                row_bytes = 8 * int(np.prod(elem_shape[1:])) * (2 if job.split else 1)
                self.rows = max(1, min(job.n_obs, ROW_SCRATCH_BYTES // max(row_bytes, 1)))
                self.ptr_t, self.idx_t, self.sbn_t = dev_in[9:12]
                scr_shape = [min(self.rows, job.n_obs)] + elem_shape[1:]
                self.scr = empty(scr_shape)
                self.scr2 = empty(scr_shape) if job.split else None
            else:
                self.rows = job.n_obs
            self.fold_cap = max(1, self.rows * self.n_strad)
            self.fold_list = empty((2 * self.fold_cap,), torch.int64) if strad.size else None
            # the folded pairs' grids and centre loads: sized by the exact host
            # count of band pairs (any one launch holds at most all of them)
            self.fold_items = job_fold_bound(job, recs, strad)
            self.fold_work = (empty((self.fold_items * dev.lib.mhs_fold_item_bytes(job.order),), torch.uint8)
                              if self.fold_items else None)
        self.far_tiles = 0
        self.tiles_fixed = False
        self.rule = rule_struct(mixed_rule)
        self.rule_tuned = False

    def _launch_struct(self, lo, hi, out_p, out2_p, stages):
        job, dev = self.job, self.dev
        return _lib.Launch(
            abi_version=_lib.ABI_VERSION, kind=self.kind, order=job.order, flags=self.flags,
            n_a=job.n_a, n_b=self.n_b_launch, n_obs=job.n_obs, n_src=job.n_src, n_cols=self.n_cols,
            obs_lo=lo, obs_hi=hi, mu=job.mu, lam=job.lam,
            obs=_ptr(self.obs_t), nrm=_ptr(self.nrm_t), rb=_ptr(self.rb_t), sb=_ptr(self.sb_t),
            src=_ptr(self.src_t), coef=_ptr(self.coef_t), gl_x=_ptr(dev.gl_x), gl_w=_ptr(dev.gl_w),
            gq=_ptr(dev.gq),
            out=out_p, out2=out2_p, strad=_ptr(self.strad_t), n_strad=self.n_strad,
            fold_list=_ptr(self.fold_list), fold_count=_ptr(self.fold_cnt), fold_err=_ptr(self.fold_err),
            fold_cap=self.fold_cap, stages=stages, far_tiles=self.far_tiles,
            fold_work=_ptr(self.fold_work), fold_work_items=self.fold_items, rule=self.rule,
            near_tiles=_ptr(self.near_tiles_t))

    def chunks(self, max_pairs: int | None = None):
        """Observer-row ranges that together cover the job, each at most
        ``max_pairs`` (observer, column) pairs (and, in the continuous layout,
        at most the element scratch's rows)."""
        job = self.job
        step = self.rows
        if max_pairs is not None:
            step = max(1, min(step, int(max_pairs) // max(self.n_cols, 1)))
        return [(lo, min(job.n_obs, lo + step)) for lo in range(0, job.n_obs, step)]

    def tune_key(self):
        return (self.job.order, self.job.n_a * self.job.n_b, bool(self.job.mixed),
                int(np.log2(max(self.n_cols, 1))))

    def _sample_blocks(self):
        """About AUTOTUNE_PAIRS pairs in three blocks of rows spread over the
        matrix (receivers near its ends and its middle see different near/far
        mixes), or all rows."""
        n = self.job.n_obs
        rows = int(min(n, max(1, AUTOTUNE_PAIRS // max(self.n_cols, 1))))
        if rows >= n and n <= self.rows:
            return [(0, n)]
        per = max(1, min(rows // 3, self.rows))     # a block fits the continuous layout's scratch
        return sorted({(max(0, min(n - per, int(f * n) - per // 2)),
                        max(0, min(n - per, int(f * n) - per // 2)) + per) for f in (1 / 6, 1 / 2, 5 / 6)})

    def rule_key(self):
        """The class a tuned mixed rule is cached under: form, whether the
        rows are the mesh's own collocation points, order, outputs, size."""
        order, outputs, _, cols = self.tune_key()
        return (self.job.kind, bool(self.job.self_rows), order, outputs, cols)

    def fix_rule(self):
        """Use the rule fixed in ``mixed_rule``, else the device's tuned rule
        for this job's class if there is one, else the library default --
        without tuning (for callers that must not synchronise)."""
        self.rule = rule_struct(mixed_rule if mixed_rule is not None else self.dev.rules.get(self.rule_key()))
        self.rule_tuned = True

    def fix_tiles(self):
        """Use the device's tuned far-pass tiles for this job's class if there
        are any, else the library default (``far_tiles`` 0), without tuning."""
        self.far_tiles = self.dev.far_tiles.get(self.tune_key(), 0) if self.contracted else 0
        self.tiles_fixed = True

    def tune_rule(self):
        """Choose the mixed mode's far-field rule (near/far switch and Gauss
        orders by distance band) under a measured error bound: on the sample
        rows the FP64 rule's passes (far, near, fold) are the reference; every
        candidate of RULE_CANDIDATES is run in the mixed mode on the same
        rows, its error (the largest over the rows of max|mixed - FP64| /
        max|FP64| in the row) measured and its passes timed (HIP events,
        the fastest of AUTOTUNE_REPEATS launches after one warm launch); of
        the candidates within RULE_ERROR_BOUND, the most accurate of those
        within RULE_TIE of the fastest wins (:func:`pick_rule`; none within
        the bound: the last, RULE_SAFE).  Cached per
        rule_key() on the device with every candidate's error and time.
        Only the on-fault interaction (``job.self_rows``) is tuned: other
        forms take the library's default, RULE_SAFE, whose error is bounded
        per pair.  A rule fixed in ``mixed_rule`` is used untuned; so is the
        library's default for the continuous layout and for jobs smaller
        than the sample (AUTOTUNE_PAIRS pairs), which a tuned rule of their
        class replaces once one exists.  Synchronises the stream."""
        if not self.job.mixed or self.rule_tuned:
            return
        self.rule_tuned = True
        dev, torch = self.dev, self.dev.torch
        if mixed_rule is not None or not autotune:
            self.rule = rule_struct(mixed_rule)
            return
        key = self.rule_key()
        if key in dev.rules:
            self.rule = rule_struct(dev.rules[key])
            return
        if not self.job.self_rows or self.cont or self.job.n_obs * self.n_cols < AUTOTUNE_PAIRS:
            self.rule = rule_struct(None)
            return
        blocks = self._sample_blocks()
        row_elems = int(np.prod(self.final.shape[1:]))
        stages = 0        # every pass: the rule also splits the fold's points
        lib = dev.lib

        def run(flags, rule, tiles):
            for lo, hi in blocks:
                off = 8 * lo * row_elems
                L = self._launch_struct(lo, hi, C.c_void_p(self.final.data_ptr() + off),
                                        C.c_void_p(self.final2.data_ptr() + off) if self.final2 is not None
                                        else None, stages)
                L.flags, L.rule, L.far_tiles = flags, rule, tiles
                _lib.check(lib.mhs_assemble(C.byref(L), self.handle), lib)

        table = {}
        with torch.cuda.stream(self.stream):
            run(self.flags & ~_lib.F_MIXED, _lib.Rule(), 0)
            ref = [self.final[lo:hi].clone() for lo, hi in blocks]
            red = tuple(range(1, self.final.dim()))
            scale = [r.abs().amax(dim=red).clamp_min(1e-300) for r in ref]
            for cand in RULE_CANDIDATES:
                rs = rule_struct(cand)
                run(self.flags, rs, 0)
                best_ms = float("inf")
                for _ in range(AUTOTUNE_REPEATS):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(self.stream)
                    run(self.flags, rs, 0)
                    e1.record(self.stream)
                    e1.synchronize()
                    best_ms = min(best_ms, e0.elapsed_time(e1))
                err = max(float(((self.final[lo:hi] - r).abs().amax(dim=red) / sc).max())
                          for (lo, hi), r, sc in zip(blocks, ref, scale))
                table[cand] = (err, best_ms)
        best = pick_rule(table, fallback=RULE_CANDIDATES[-1])
        dev.rules[key] = best
        dev.rule_table[key] = {_rule_name(c): {"relmax": e, "ms": t} for c, (e, t) in table.items()}
        self.rule = rule_struct(best)

    def tune(self):
        """Pick the far pass's column tiles per block: each candidate is timed
        (HIP events, after one warm launch, the faster of AUTOTUNE_REPEATS
        launches) on the far pass over about AUTOTUNE_PAIRS pairs in three
        blocks of rows spread over the matrix; the fastest is cached per
        (order, outputs, precision, size class) on the device, with the
        timings.  Only the grouping of pairs into blocks changes, never a
        value; the sample rows are rewritten by the real launch.
        Synchronises the stream once per candidate of a new key."""
        if not self.contracted or self.far_tiles:
            return self.far_tiles
        self.tune_rule()
        key = self.tune_key()
        dev, torch = self.dev, self.dev.torch
        if key in dev.far_tiles:
            self.far_tiles = dev.far_tiles[key]
            return self.far_tiles
        if not autotune:
            return 0
        blocks = self._sample_blocks()
        times = {}
        with torch.cuda.stream(self.stream):
            for ct in tile_candidates(self.job.mixed):
                self.far_tiles = ct
                for lo, hi in blocks:
                    self.launch_rows(lo, hi, _lib.STAGE_Q)
                best_ms = float("inf")
                for _ in range(AUTOTUNE_REPEATS):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(self.stream)
                    for lo, hi in blocks:
                        self.launch_rows(lo, hi, _lib.STAGE_Q)
                    e1.record(self.stream)
                    e1.synchronize()
                    best_ms = min(best_ms, e0.elapsed_time(e1))
                times[ct] = best_ms
        # the sample separates the candidates by a few percent at most, less
        # than the full launch's spread, so near-ties go to the most tiles per
        # block (4 tiles per block sampled as fast as 32 and ran the full
        # 10^4 x 10^4 far pass 6-8 % slower)
        fastest = min(times.values())
        best = max(ct for ct, t in times.items() if t <= fastest * (1.0 + AUTOTUNE_TIE))
        dev.far_tiles[key] = best
        dev.tune_times[key] = times
        self.far_tiles = best
        return best

    def _ready(self):
        """Tune (or fix) the launch shape and the mixed rule before the first
        launch (the autotuner synchronises on its first visit of a shape)."""
        if self.contracted and not self.far_tiles and not self.tiles_fixed:
            self.tune()
        if self.job.mixed and not self.rule_tuned:
            self.tune_rule()

    def descriptor(self, lo: int, hi: int, stages: int = 0):
        """The launch descriptor (``_lib.Launch``) of observer rows [lo, hi)
        (at most ``self.rows`` of them in the continuous layout)."""
        job = self.job
        if self.cont:
            if hi - lo > self.rows:
                raise ValueError(f"rows [{lo}, {hi}) exceed the element scratch ({self.rows} rows)")
            out_p, out2_p = _ptr(self.scr), _ptr(self.scr2)
        else:
            row_elems = int(np.prod(self.final.shape[1:]))
            out_p = C.c_void_p(self.final.data_ptr() + 8 * lo * row_elems)
            out2_p = (C.c_void_p(self.final2.data_ptr() + 8 * lo * row_elems)
                      if job.split else None)
        return self._launch_struct(lo, hi, out_p, out2_p, stages)

    def epilogue(self, lo: int, hi: int, stages: int = 0):
        """What follows the assembly of rows [lo, hi) on the stream: the
        continuous layout's gather, the elastic stress as total minus
        eigenstress, the strain."""
        job, lib = self.job, self.dev.lib
        if self.cont:
            self._gather(lo, hi)
        if job.split and job.subtract and (stages == 0 or stages == _ALL_STAGES):
            # elastic = finished total - finished eigenstress, as the reference forms it
            row_elems = int(np.prod(self.final.shape[1:]))
            off = 8 * lo * row_elems
            _lib.check(lib.mhs_subtract(C.c_void_p(self.final.data_ptr() + off),
                                        C.c_void_p(self.final2.data_ptr() + off),
                                        (hi - lo) * row_elems, self.handle), lib)
        if job.strain:
            row_elems = int(np.prod(self.final.shape[1:]))
            _lib.check(lib.mhs_strain(C.c_void_p(self.final.data_ptr() + 8 * lo * row_elems), hi - lo,
                                      job.n_dof, job.mu, job.lam, self.handle), lib)

    def launch_rows(self, lo: int, hi: int, stages: int = 0):
        """Enqueue every pass for observer rows [lo, hi) (at most ``self.rows``
        of them in the continuous layout); no host synchronisation (except
        the autotuner's first visit of a new launch shape)."""
        lib = self.dev.lib
        self._ready()
        L = self.descriptor(lo, hi, stages)
        _lib.check(lib.mhs_assemble(C.byref(L), self.handle), lib)
        self.epilogue(lo, hi, stages)

    def launch(self, stages: int = 0):
        """Enqueue every pass on the stream (no host synchronisation)."""
        for lo, hi in self.chunks():
            self.launch_rows(lo, hi, stages)

    def graph(self, stages: int = 0):
        """Capture the whole launch sequence as one HIP graph on this job's
        stream (which must not be the default stream); ``replay()`` re-runs
        the assembly into the same buffers.  The sequence has no memset node
        (counters are cleared by a fill kernel) and no host synchronisation;
        the far pass's launch shape is tuned before the capture."""
        torch = self.dev.torch
        if self.stream == torch.cuda.default_stream(self.dev.dev):
            raise ValueError("graph capture needs a Prepared on a non-default stream")
        if self.contracted and not self.far_tiles:
            self.tune()
            torch.cuda.synchronize()
        if self.job.mixed and not self.rule_tuned:
            self.tune_rule()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self.launch(stages)
        return g

    def _gather(self, lo, hi):
        job, lib = self.job, self.dev.lib
        for s_t, f_t in ((self.scr, self.final), (self.scr2, self.final2)):
            if s_t is None:
                continue
            dst = C.c_void_p(f_t.data_ptr() + 8 * lo * int(np.prod(f_t.shape[1:])))
            if job.kind == "interaction":
                rc = lib.mhs_gather(_ptr(s_t), dst, hi - lo, self.n_cols, job.n_nodes, 0,
                                    _ptr(self.ptr_t), _ptr(self.idx_t), _ptr(self.sbn_t),
                                    job.n_a, job.n_b, self.handle)
            else:
                lead = (hi - lo) * int(np.prod(s_t.shape[1:-2]))
                rc = lib.mhs_gather(_ptr(s_t), dst, lead, self.n_cols, job.n_nodes, 3,
                                    _ptr(self.ptr_t), _ptr(self.idx_t), None, 0, 0, self.handle)
            _lib.check(rc, lib)

    HIST_KEYS = ("pairs", "direct_near", "direct_far12", "direct_far8", "image_near", "image_far12",
                 "image_far8", "fold_pairs", "q_points", "q_pairs", "mirror_near", "mirror_far12",
                 "mirror_far8", "m_pairs", "m_direct_near", "m_direct_far", "m_direct_points",
                 "m_image_near", "m_image_far", "m_image_points", "m_q_pairs", "m_q_points",
                 "m_q_chunks", "m_image_chunks")

    def histogram(self) -> dict:
        """Branch counts of every pair, by the kernels' own classification
        (the m_* counts: the mixed rule's far pass, per (row, source) pair)."""
        torch, lib = self.dev.torch, self.dev.lib
        with torch.cuda.stream(self.stream):
            cnt = torch.zeros(len(self.HIST_KEYS), dtype=torch.int64, device=self.dev.dev)
            L = self._launch_struct(0, self.job.n_obs, None, None, 0)
            _lib.check(lib.mhs_histogram(C.byref(L), _ptr(cnt), self.handle), lib)
            vals = cnt.cpu().tolist()
        return dict(zip(self.HIST_KEYS, vals))

    FOLD_HIST_KEYS = ("fold_near_points", "fold_far_points", "fold_far_gauss_points", "fold_items")

    def fold_histogram(self) -> dict:
        """Point counts of the fold's passes from the last assembly's fold
        buffers (the rows of its last row chunk; the whole job when it ran
        as one launch), by the kernels' own classification; syncs."""
        torch, lib = self.dev.torch, self.dev.lib
        with torch.cuda.stream(self.stream):
            cnt = torch.zeros(len(self.FOLD_HIST_KEYS), dtype=torch.int64, device=self.dev.dev)
            if self.n_strad:
                lo, hi = self.chunks()[-1]
                L = self.descriptor(lo, hi)
                _lib.check(lib.mhs_fold_histogram(C.byref(L), _ptr(cnt), self.handle), lib)
            vals = cnt.cpu().tolist()
        return dict(zip(self.FOLD_HIST_KEYS, vals))

    def check_fold(self):
        """Raise if a folded pair overflowed its grid or the pair list (syncs)."""
        if self.n_strad:
            err = int(self.fold_err.cpu()[0])
            if err:
                raise RuntimeError(f"mhs_amd: fold grid/list/work overflow (flags {err})")


class Batch:
    """Row ranges of several prepared jobs enqueued as one batch on one
    stream (``mhs_assemble_batch``): each tiled pass of every range goes
    into a device work queue drained by one persistent launch per kernel
    instance, then each range's fold and epilogue follow.  Results are
    bitwise those of ``Prepared.launch_rows`` on each range.  A job may
    appear once per batch (its fold buffers and continuous-layout scratch
    serve one row range at a time); batches on one stream run in order.
    The workspaces (pinned host, device) are held until :meth:`done`."""

    def __init__(self, items, stream, dev: Device | None = None):
        dev = dev or device()
        self.dev, self.stream, self.items = dev, stream, list(items)
        torch, lib = dev.torch, dev.lib
        ids = [id(p) for p, _, _ in self.items]
        if len(set(ids)) != len(ids):
            raise ValueError("a job may appear once per batch")
        for prep, _, _ in self.items:
            prep._ready()
        n = len(self.items)
        arr = (_lib.Launch * max(n, 1))()
        for i, (prep, lo, hi) in enumerate(self.items):
            arr[i] = prep.descriptor(lo, hi)
        self._arr = arr
        nbytes = int(lib.mhs_batch_workspace_bytes(arr, n))
        if nbytes < 0:
            _lib.check(-1, lib)
        nbytes = max(nbytes, 256)
        # torch's pinned and device blocks are at least 512-byte aligned
        self.host_ws = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        handle = C.c_void_p(stream.cuda_stream)
        with torch.cuda.stream(stream):
            self.dev_ws = torch.empty(nbytes, dtype=torch.uint8, device=dev.dev)
            _lib.check(lib.mhs_assemble_batch(arr, n, C.c_void_p(self.dev_ws.data_ptr()),
                                              C.c_void_p(self.host_ws.data_ptr()), nbytes, handle), lib)
            for prep, lo, hi in self.items:
                if prep.stream != stream:
                    raise ValueError("every job of a batch must be prepared on the batch's stream")
                prep.epilogue(lo, hi)
            self.event = torch.cuda.Event()
            self.event.record(stream)
        self.nbytes = nbytes

    def query(self) -> bool:
        return self.event.query()

    def done(self) -> None:
        """Wait for the batch and release its workspaces."""
        self.event.synchronize()
        self.host_ws = self.dev_ws = None


def fold_pairs_bound(obs: np.ndarray, recs: np.ndarray, strad: np.ndarray, chunk: int = 4096) -> int:
    """An upper bound on the (observer, straddling source) pairs in the fold
    band (fold.py:84-118: receiver shallower than 10 eps, within 60 eps of the
    crease centre): the host count plus a margin for the pairs a last-ulp
    difference could move across the band's edge."""
    if strad.size == 0:
        return 0
    eps = recs[strad, geometry.S_EPS]
    cx = recs[strad, geometry.S_CREASE]
    cy = recs[strad, geometry.S_CREASE + 1]
    n = 0
    for lo in range(0, obs.shape[0], chunk):
        o = obs[lo:lo + chunk]
        depth_ok = (-o[:, 2:3]) < 10.0 * eps[None, :] * (1.0 + 1e-12)
        reach = np.hypot(o[:, 0:1] - cx[None, :], o[:, 1:2] - cy[None, :])
        n += int(np.count_nonzero(depth_ok & (reach < 60.0 * eps[None, :] * (1.0 + 1e-12))))
    return n + 64


def job_records(job: Job) -> np.ndarray:
    """The job's source records (geometry.source_records), computed once and
    kept on the job for the arrays they were computed from."""
    hit = job.extra.get("records")
    if hit is None or hit[0] is not job.tris or hit[1] is not job.eps:
        hit = job.extra["records"] = (job.tris, job.eps, geometry.source_records(job.tris, job.eps))
    return hit[2]


def job_fold_bound(job: Job, recs: np.ndarray, strad: np.ndarray) -> int:
    """fold_pairs_bound of the job, computed once and kept on the job for
    the observer and source arrays it was computed from."""
    if strad.size == 0:
        return 0
    hit = job.extra.get("fold_bound")
    if hit is None or hit[0] is not job.obs or hit[1] is not job.tris or hit[2] is not job.eps:
        hit = job.extra["fold_bound"] = (job.obs, job.tris, job.eps, fold_pairs_bound(job.obs, recs, strad))
    return hit[3]


def job_bytes(job: Job) -> int:
    """Device bytes a :class:`Prepared` of ``job`` holds: outputs, the
    continuous layout's element scratch, the fold list and the inputs."""
    out = 8 * int(np.prod(job.out_shape())) * (2 if job.split else 1)
    n_cols = job.n_src * job.K
    scratch = 0
    rows = job.n_obs
    if job.ids is not None:
        shape = list(job.out_shape(n_cols))
        if job.kind == "interaction":
            shape[-1] = 3
        row_bytes = 8 * int(np.prod(shape[1:])) * (2 if job.split else 1)
        rows = max(1, min(job.n_obs, ROW_SCRATCH_BYTES // max(row_bytes, 1)))
        scratch = rows * row_bytes
    n_strad = int(np.sum(np.asarray(job.tris)[:, :, 2].max(axis=1) > -3.0 * np.asarray(job.eps)))
    fold = 16 * max(1, rows * n_strad) if job.fold and n_strad else 0
    if fold:
        recs = job_records(job)
        strad = np.flatnonzero(recs[:, geometry.S_STRADDLE] != 0.0)
        fold += FOLD_ITEM_BYTES * job_fold_bound(job, recs, strad)
    inputs = 8 * (job.n_obs * 3 * (2 + job.n_a) + job.n_src * (geometry.SRC_STRIDE + 3 * max(job.n_b, 0)))
    return int(out + scratch + fold + inputs)


def run(job: Job, stream=None, dev: Device | None = None):
    """Execute ``job`` on the GPU; returns (out, out2) device tensors in the
    final layout (out2 is the eigenstress part for parts=True, else None)."""
    prep = Prepared(job, stream, dev)
    with prep.dev.torch.cuda.stream(prep.stream):
        prep.launch()
    prep.check_fold()
    return prep.final, prep.final2


#: Device results larger than this go to the host through pinned staging.
STAGED_COPY_BYTES = 16 << 20
#: Bytes per pinned staging buffer (two per process, reused).
STAGE_BYTES = 64 << 20
_STAGING: dict = {}
_STAGING_LOCK = threading.Lock()


def to_host(t, out=None):
    """Copy a device tensor into ``out`` (float64 or float32 ndarray) or a
    new float64 array.  Large results go through two pinned staging buffers
    on a copy stream: the device-to-host copy of one chunk overlaps the host
    copy of the previous one into ``out`` (a pageable-memory copy of the
    whole tensor runs several times slower)."""
    n = int(t.numel())
    if out is None:
        out = np.empty(tuple(t.shape), dtype=np.float64)
    if n * 8 < STAGED_COPY_BYTES or not out.flags.c_contiguous:
        if out.dtype == np.float32:
            out[...] = t.float().cpu().numpy()
        else:
            out[...] = t.cpu().numpy()
        return out
    with _STAGING_LOCK:
        return _staged_copy(t, out, n)


def _staged_copy(t, out, n):
    torch = _torch()
    dev = t.device
    key = (dev.index, STAGE_BYTES)
    if key not in _STAGING:
        ch = STAGE_BYTES // 8
        _STAGING[key] = ([torch.empty(ch, dtype=torch.float64, pin_memory=True) for _ in range(2)],
                         torch.cuda.Stream(device=dev))
    bufs, copy = _STAGING[key]
    ch = int(bufs[0].numel())
    src = t.reshape(-1)
    dst = out.reshape(-1)
    copy.wait_stream(torch.cuda.current_stream(dev))
    pending = None
    for i, lo in enumerate(range(0, n, ch)):
        hi = min(n, lo + ch)
        b = bufs[i % 2]
        with torch.cuda.stream(copy):
            b[: hi - lo].copy_(src[lo:hi], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(copy)
        if pending is not None:
            plo, phi, pb, pev = pending
            pev.synchronize()
            dst[plo:phi] = pb[: phi - plo].numpy()
        pending = (lo, hi, b, ev)
    plo, phi, pb, pev = pending
    pev.synchronize()
    dst[plo:phi] = pb[: phi - plo].numpy()
    return out
