"""Device execution of one assembly job: buffers, launches, gather, readout.

PyTorch-ROCm supplies device memory and the stream; every arithmetic step on
the pair path is a HIP kernel behind the C-ABI (``_lib``).  There is no CPU
fallback: without a GPU or without the built library this module raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib, geometry

ROW_SCRATCH_BYTES = 4 << 30      # continuous-layout element-major scratch per launch


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
    nrm: np.ndarray | None = None      # (n_obs, 3)
    rb: np.ndarray | None = None       # (n_obs, n_a, 3)
    sb: np.ndarray | None = None       # (n_src | n_nodes, n_b, 3)
    ids: np.ndarray | None = None      # (n_src, K) continuous columns
    n_nodes: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def K(self) -> int:
        return (self.order + 1) * (self.order + 2) // 2

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
        self._coef = {}

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


_DEVICE = None


def device() -> Device:
    global _DEVICE
    if _DEVICE is None:
        _DEVICE = Device(0)
    return _DEVICE


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _csr(ids: np.ndarray, n_nodes: int):
    """Element columns of every shared node, ascending: (ptr, idx)."""
    flat = ids.reshape(-1)
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=n_nodes)
    ptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    return ptr, order.astype(np.int64)


def run(job: Job, stream=None, dev: Device | None = None):
    """Execute ``job`` on the GPU; returns (out, out2) device tensors in the
    final layout (out2 is the eigenstress part for parts=True, else None)."""
    dev = dev or device()
    torch, lib = dev.torch, dev.lib
    st = stream if stream is not None else torch.cuda.current_stream(dev.dev)
    handle = C.c_void_p(st.cuda_stream)
    K = job.K
    n_cols = job.n_src * K
    recs = geometry.source_records(job.tris, job.eps)
    strad = np.flatnonzero(recs[:, geometry.S_STRADDLE] != 0.0).astype(np.int64) if job.fold else np.zeros(0, np.int64)
    with torch.cuda.stream(st):
        obs_t = dev.put(job.obs)
        src_t = dev.put(recs)
        nrm_t = dev.put(job.nrm)
        rb_t = dev.put(job.rb)
        cont = job.ids is not None
        sb_t = None if cont else dev.put(job.sb)
        strad_t = dev.put(strad, np.int64) if strad.size else None
        coef_t = dev.coef(job.mu, job.lam)
        final = torch.empty(job.out_shape(), dtype=torch.float64, device=dev.dev)
        final2 = torch.empty_like(final) if job.parts else None
        kind = _lib.KIND[job.kind]
        flags = ((_lib.F_FOLD if job.fold else 0) | (_lib.F_SUBTRACT if job.subtract else 0)
                 | (_lib.F_PARTS if job.parts else 0))
        n_b_launch = 3 if (cont and job.kind == "interaction") else job.n_b
        if cont:
            elem_shape = list(job.out_shape(n_cols))
            if job.kind == "interaction":
                elem_shape[-1] = 3
            row_bytes = 8 * int(np.prod(elem_shape[1:])) * (2 if job.parts else 1)
            rows = max(1, min(job.n_obs, ROW_SCRATCH_BYTES // max(row_bytes, 1)))
            ptr, idx = _csr(job.ids, job.n_nodes)
            ptr_t, idx_t = dev.put(ptr, np.int64), dev.put(idx, np.int64)
            sbn_t = dev.put(job.sb) if job.kind == "interaction" else None
        else:
            rows = job.n_obs
        fold_cap = max(1, rows * len(strad))
        fold_list = torch.empty(2 * fold_cap, dtype=torch.int64, device=dev.dev) if strad.size else None
        fold_cnt = torch.zeros(2, dtype=torch.int32, device=dev.dev)
        fold_err = torch.zeros(1, dtype=torch.int32, device=dev.dev)
        for lo in range(0, job.n_obs, rows):
            hi = min(job.n_obs, lo + rows)
            if cont:
                scr_shape = [hi - lo] + elem_shape[1:]
                scr = torch.empty(scr_shape, dtype=torch.float64, device=dev.dev)
                scr2 = torch.empty_like(scr) if job.parts else None
                out_p, out2_p = _ptr(scr), _ptr(scr2)
            else:
                row_elems = int(np.prod(final.shape[1:]))
                out_p = C.c_void_p(final.data_ptr() + 8 * lo * row_elems)
                out2_p = C.c_void_p(final2.data_ptr() + 8 * lo * row_elems) if job.parts else None
            L = _lib.Launch(
                abi_version=_lib.ABI_VERSION, kind=kind, order=job.order, flags=flags,
                n_a=job.n_a, n_b=n_b_launch, n_obs=job.n_obs, n_src=job.n_src, n_cols=n_cols,
                obs_lo=lo, obs_hi=hi, mu=job.mu, lam=job.lam,
                obs=_ptr(obs_t), nrm=_ptr(nrm_t), rb=_ptr(rb_t), sb=_ptr(sb_t), src=_ptr(src_t),
                coef=_ptr(coef_t), gl_x=_ptr(dev.gl_x), gl_w=_ptr(dev.gl_w),
                out=out_p, out2=out2_p, strad=_ptr(strad_t), n_strad=len(strad),
                fold_list=_ptr(fold_list), fold_count=_ptr(fold_cnt), fold_err=_ptr(fold_err),
                fold_cap=fold_cap)
            _lib.check(lib.mhs_assemble(C.byref(L), handle), lib)
            if cont:
                for s_t, f_t in ((scr, final), (scr2, final2)):
                    if s_t is None:
                        continue
                    dst = C.c_void_p(f_t.data_ptr() + 8 * lo * int(np.prod(f_t.shape[1:])))
                    if job.kind == "interaction":
                        rc = lib.mhs_gather(_ptr(s_t), dst, hi - lo, n_cols, job.n_nodes, 0, _ptr(ptr_t),
                                            _ptr(idx_t), _ptr(sbn_t), job.n_a, job.n_b, handle)
                    else:
                        lead = (hi - lo) * int(np.prod(s_t.shape[1:-2]))
                        rc = lib.mhs_gather(_ptr(s_t), dst, lead, n_cols, job.n_nodes, 3, _ptr(ptr_t),
                                            _ptr(idx_t), None, 0, 0, handle)
                    _lib.check(rc, lib)
        if job.strain:
            _lib.check(lib.mhs_strain(_ptr(final), job.n_obs, job.n_dof, job.mu, job.lam, handle), lib)
        err = int(fold_err.cpu()[0]) if strad.size else 0
    if err:
        raise RuntimeError(f"mhs_amd: fold grid/list overflow (flags {err}); the polar grid of a "
                           f"folded pair exceeded its panel budget")
    return final, final2


def to_host(t, out=None):
    """Copy a device tensor into ``out`` (float64 or float32 ndarray) or a new array."""
    if out is None:
        return t.cpu().numpy()
    if out.dtype == np.float32:
        out[...] = t.to(dtype=t.new_empty(0).float().dtype).cpu().numpy()
    else:
        out[...] = t.cpu().numpy()
    return out
