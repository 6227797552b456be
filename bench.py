"""Benchmark: interaction matrix, n = 10,000 strike/strike, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode mixed|fp64]
                    [--cpu-cols C] [--no-cpu]

A step is one full assembly of the workload with every input already resident
in HBM and the (10^4, 10^4, 1, 1) output left in HBM: the pair pass (direct +
image R-family / folded mirror + eigenstress), the Q-family pass and the
Boussinesq fold pass.  ``value`` is pair-evals/s of the requested mode (mixed
by default: far-field quadrature points in FP32, FP64 accumulation); the other
mode is timed beside it and the mixed result's max error relative to the FP64
result is measured on all 10^8 entries.  Per-pass kernel times come from HIP
events recorded on the launch stream; the roofline of the dominant kernel uses
the kernels' own branch histogram and the static FLOP model of
mhs_amd/flops.py.  The CPU baseline is the FP64 oracle (a numpy restatement
of the reference) on every receiver against a stratified seeded sample of
source columns on the host's CPU share (16 processes), extrapolated to the
whole matrix per stratum.  The served stream (512 requests at a frozen rate,
both precisions) follows.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair-evals/s, interaction n=10k, mixed vs FP64 @1e-6 rel; served p50/p99"
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}     # MI355X vector FP64, FP32 (published)
CPU_SEED = 20261203


#: The reference's published single-core figure (README "Across cores"): the
#: strike/strike interaction of an 800-element fault, 640k pairs in 30.1 s.
PUBLISHED_US_PER_PAIR = 47.1
#: Worker processes for the CPU baseline: the host's CPU share for one GPU.
CPU_SHARE = 16


def _cpu_block(task):
    """One task of the CPU baseline: ``(workload, columns)``; returns
    (pairs, CPU seconds of this process for the task)."""
    from mhs_amd import workloads
    from oracle import api as oapi
    what, cols = task
    t0 = time.process_time()
    if what == "c3":
        tris, eps = workloads.c3_interaction()
    else:           # the published benchmark's size: a buried 800-element plane
        tris, eps = workloads.orient_up(workloads.grid_mesh(workloads.plane(40.0, 20.0, 20, 20, 60.0, -2.0))), 0.1
    K = oapi.interaction_matrix(tris[cols], 30.0, 30.0, eps, receiver="strike", source="strike",
                                obs_tris=tris)
    return int(K.shape[0] * K.shape[1]), time.process_time() - t0


def _pool_run(tasks, procs):
    """Run tasks on a pool of ``procs`` spawned processes with BLAS pinned to
    one thread; returns ([(pairs, cpu_s)] in task order, wall seconds)."""
    import multiprocessing as mp
    env = {"OMP_NUM_THREADS": "1", "OPENBLAS_NUM_THREADS": "1", "MKL_NUM_THREADS": "1"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        with mp.get_context("spawn").Pool(procs) as pool:
            t0 = time.perf_counter()
            res = pool.map(_cpu_block, tasks, chunksize=1)
            return res, time.perf_counter() - t0
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _physical_cores():
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        return len({ln for ln in out.splitlines() if ln and not ln.startswith("#")}) or None
    except (OSError, subprocess.SubprocessError):
        return None


def cpu_baseline(n_cols: int) -> dict:
    """The FP64 oracle (a numpy restatement of the reference, which runs one
    source column at a time vectorised over the receivers) on all 10^4
    receivers x ``n_cols`` seeded source columns of the workload, one column
    per task on CPU_SHARE spawned processes (BLAS pinned to one thread),
    surface-reaching (folded, costliest) columns first.  The sample is
    stratified (surface-reaching sources, shallowest vertex above 3 eps
    depth, and the rest, in proportion to their counts) and each task's CPU
    seconds are kept per stratum; ``value`` is the whole 10^4 x 10^4 matrix
    extrapolated from them, spread evenly over the processes (the sample's
    own wall-clock rate, which two long folded columns dominate, is
    reported beside it).
    Also: a one-process run (one column per stratum) and the restatement's
    serial us/pair on the published benchmark's size beside the published
    figure."""
    from mhs_amd import workloads
    rng = np.random.default_rng(CPU_SEED)
    tris, eps = workloads.c3_interaction()
    n = tris.shape[0]
    shallow = tris[:, :, 2].max(axis=1) > -3.0 * eps
    strad, rest = np.flatnonzero(shallow), np.flatnonzero(~shallow)
    n_sh = int(round(n_cols * len(strad) / n))
    cols = np.concatenate([rng.choice(strad, size=n_sh, replace=False),
                           rng.choice(rest, size=n_cols - n_sh, replace=False)])
    procs = min(CPU_SHARE, len(cols))
    res, wall = _pool_run([("c3", np.array([c])) for c in cols], procs)
    pairs = sum(p for p, _ in res)
    per = {}
    for name, sel, count in (("surface_reaching", slice(0, n_sh), len(strad)),
                             ("buried", slice(n_sh, None), len(rest))):
        p = sum(x for x, _ in res[sel])
        c = sum(t for _, t in res[sel])
        per[name] = {"columns_sampled": len(res[sel]), "columns_in_matrix": int(count),
                     "cpu_us_per_pair": 1e6 * c / max(p, 1)}
    total_cpu = sum(v["columns_in_matrix"] * n * v["cpu_us_per_pair"] * 1e-6 for v in per.values())
    one, one_wall = _pool_run([("c3", np.array([cols[0]])), ("c3", np.array([cols[-1]]))], 1)
    pub, _ = _pool_run([("plane800", np.arange(20))], 1)
    phys = _physical_cores()
    return {"value": n * n / (total_cpu / procs), "unit": "pair-evals/s", "cores": procs, "kind": "port",
            "sample": f"all 10,000 receivers x {n_cols} seeded source columns (seed {CPU_SEED}; {n_sh} of them "
                      f"surface-reaching, as in the whole mesh) of the same workload, FP64 oracle, "
                      f"{pairs:.3e} pairs (their fold pairs included) in {wall:.1f} s wall on {procs} "
                      f"processes (the host's CPU share for one GPU); value = the whole 10^4 x 10^4 "
                      f"matrix extrapolated from each stratum's CPU seconds per pair, spread evenly "
                      f"over the {procs} processes",
            "sample_wall_pair_evals_per_s": pairs / wall,
            "host": {"os_cpu_count": os.cpu_count(), "lscpu_physical_cores": phys},
            "strata": per,
            "one_process": {"columns": 2, "pairs": sum(p for p, _ in one), "wall_s": one_wall,
                            "us_per_pair": 1e6 * one_wall / sum(p for p, _ in one)},
            "extrapolated": {"label": "extrapolated from the per-stratum CPU seconds, not measured",
                             "full_matrix_cpu_s": total_cpu,
                             "full_matrix_wall_s_on_cores": total_cpu / procs,
                             "pair_evals_per_s_on_cores": n * n / (total_cpu / procs),
                             "pair_evals_per_s_on_physical_cores": (n * n / (total_cpu / phys)) if phys else None},
            "vs_published": {"published_us_per_pair": PUBLISHED_US_PER_PAIR,
                             "port_us_per_pair": 1e6 * pub[0][1] / pub[0][0],
                             "case": "strike/strike interaction, buried 800-element plane, 20 source columns, "
                                     "one process (the published case: 800 elements, one core)"}}


def _traffic(mode: str):
    """HBM bytes per far-pass launch from the committed counter runs
    (FETCH_SIZE, WRITE_SIZE, each scaled by a copy / fill of known size), or
    None when they were taken in the other mode or are absent."""
    path = os.path.join(ROOT, "profiles", "k_far_mixed_traffic.json")
    if mode != "mixed" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["hbm_bytes_per_launch"]


def counter_pass(mode: str, workload: str = "c3") -> None:
    """The program that hardware counter runs profile: the workload's tuners
    (column tiles; in the mixed mode also the rule), then one whole step
    (far, near and fold passes), then a 256 MiB fill and a 256 MiB device
    copy -- the known-size calibration of FETCH_SIZE / WRITE_SIZE.  The last
    dispatch of each kernel before the fill is the full-size one.  Prints one
    JSON line: the modelled FLOPs of that step per kernel (the counted
    FP32 / FP64 operations are compared against it), the rule and tiles."""
    import torch
    from mhs_amd import flops, runtime
    mixed = mode == "mixed"
    prep = runtime.Prepared(_job(mixed, workload))
    with torch.cuda.stream(prep.stream):
        prep.launch()                   # the tuners run first
    torch.cuda.synchronize()
    prep.check_fold()
    h, fh = prep.histogram(), prep.fold_histogram()
    ff = flops.flops_fold(fh, mixed)
    if mixed:
        fm = flops.flops_mixed(h)
        model = {"far": {"fp32": fm["far_fp32"], "fp64": fm["far_fp64"]}, "near": {"fp32": 0.0, "fp64": fm["near"]}}
    else:
        fl = flops.flops(h)
        model = {"far": {"fp32": 0.0, "fp64": fl["far"]}, "near": {"fp32": 0.0, "fp64": fl["near"]}}
    model["fold_near_points"] = {"fp32": 0.0, "fp64": ff["near"]}
    model["fold_far_points"] = {"fp32": ff["far_fp32"], "fp64": ff["far_fp64"]}
    dev = runtime.device()
    key = prep.rule_key()
    rule = runtime._rule_name(dev.rules[key]) if mixed and key in dev.rules else None
    a = torch.empty(1 << 25, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"counter_pass": {"mode": mode, "workload": workload, "far_tiles": prep.far_tiles, "rule": rule,
                                       "model_flop": model, "histogram": h, "fold_histogram": fh}}), flush=True)


def source_major_timing() -> list:
    """HIP-event time of the source-major copy (mhs_source_major) of
    interaction-shaped device matrices: (receivers, sources, n_a, n_b) ->
    (sources, n_b, n_a, receivers); bytes = 8 read + 8 written per entry."""
    import ctypes as C
    import torch
    from mhs_amd import _lib
    lib = _lib.load()
    rows = []
    for (nr, ns, na, nb) in ((10000, 10000, 1, 1), (2000, 2000, 3, 3), (12000, 12000, 3, 2)):
        a = torch.randn(nr, ns, na, nb, dtype=torch.float64, device="cuda")
        b = torch.empty(ns, nb, na, nr, dtype=torch.float64, device="cuda")
        st = torch.cuda.current_stream()
        h = C.c_void_p(st.cuda_stream)
        args = (C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), nr, ns, na, nb, h)
        _lib.check(lib.mhs_source_major(*args), lib)
        torch.cuda.synchronize()
        if not torch.equal(b, a.permute(1, 3, 2, 0)):
            raise RuntimeError("source-major copy differs from the permutation")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(10):
            _lib.check(lib.mhs_source_major(*args), lib)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / 10
        rows.append({"shape": [nr, ns, na, nb], "ms": ms, "GB/s": 16 * a.numel() / ms / 1e6})
        del a, b
    return rows


def shared_nodes_timing() -> list:
    """Wall time of the continuous layout's shared-node table, host
    (numpy, mhs_amd.shared.shared_nodes) against device
    (shared_nodes_device, host-to-device and back included), on the
    six-thousand-triangle three-fault network and on single dipping planes
    of 40,000 and 490,000 triangles; the tables are checked equal."""
    from mhs_amd import workloads
    from mhs_amd.shared import shared_nodes, shared_nodes_device
    cases = [("three-fault network, grouped",) + workloads.c4_network()[:2]]
    for n in (200, 700):
        P = workloads.plane(100.0, 50.0, n, n // 2, 60.0, -1.0)
        cases.append((f"plane {n} x {n // 2} cells", workloads.orient_up(workloads.grid_mesh(P)), None))
    rows = []
    for name, tris, groups in cases:
        t0 = time.perf_counter()
        host = shared_nodes(tris, groups=groups)
        t_host = time.perf_counter() - t0
        shared_nodes_device(tris, groups=groups)              # warm: library, allocator
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            dev = shared_nodes_device(tris, groups=groups)
            times.append(time.perf_counter() - t0)
        if not (np.array_equal(dev.ids, host.ids) and np.array_equal(dev.boundary, host.boundary)):
            raise RuntimeError(f"device shared-node table differs from the host's on {name}")
        rows.append({"mesh": name, "n_tri": int(tris.shape[0]), "nodes": int(host.n),
                     "host_s": t_host, "device_s": float(np.median(times)),
                     "speedup": t_host / float(np.median(times))})
    return rows


def _time_job(job, steps: int):
    """Wall ms per assembly of a prepared job (one warm assembly first, its
    rule and launch shape tuned), and the device result."""
    import torch
    from mhs_amd import runtime
    prep = runtime.Prepared(job)
    with torch.cuda.stream(prep.stream):
        prep.launch()
        torch.cuda.synchronize()
        prep.check_fold()
        t0 = time.perf_counter()
        for _ in range(steps):
            prep.launch()
        torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps, prep


def _modes(plan, pairs: int, steps: int) -> dict:
    """FP64 and mixed timings of one workload and the mixed result's relmax
    against FP64 (max |mixed - FP64| / max |FP64| over every entry)."""
    out, res = {}, {}
    for mode in ("fp64", "mixed"):
        ms, prep = _time_job(plan(mode), steps)
        out[mode] = {"ms": ms, "pair_evals_per_s": pairs / (ms / 1e3), "shape": list(prep.final.shape)}
        res[mode] = prep.final.clone()
        del prep
    d = (res["mixed"] - res["fp64"]).abs().max().item() / res["fp64"].abs().max().item()
    out["mixed_relmax_vs_fp64"] = d
    out["mixed_over_fp64"] = out["fp64"]["ms"] / out["mixed"]["ms"]
    out["output_bytes"] = int(res["fp64"].numel() * 8)
    return out


def c4_timing(steps: int = 2) -> dict:
    """The three-fault network's continuous-layout interaction: about 6,000
    triangles welded per fault into shared nodes, order 2, per-source eps,
    the default receiver (strike, dip, normal) and source (strike, dip):
    (nodes, nodes, 3, 2) FP64 left in HBM.  Wall ms per assembly in FP64 and
    in the mixed mode (one warm assembly first), pair-evals/s counted as
    (receiver node, source element node column) pairs, and the mixed
    result's relmax against FP64."""
    import mhs_amd as M
    from mhs_amd import api, workloads
    tris, groups, eps = workloads.c4_network()
    sn = M.shared_nodes(tris, groups=groups)
    mat = M.Material(30.0, 30.0)

    def plan(mode):
        with M.precision(mode):
            return api.plan(M.interaction_matrix, tris, mat, eps, order=2, layout="continuous", shared=sn,
                            source=("strike", "dip"))
    out = {"n_tri": int(tris.shape[0]), "nodes": int(sn.n)}
    out.update(_modes(plan, int(sn.n) * int(tris.shape[0]) * 6, steps))
    return out


def c2_timing(steps: int = 3) -> dict:
    """The 2,000-triangle Gaussian-field fault: stress_matrix at 4,000
    observers (4000, 3, 3, 2000, 3) and the default 3 x 3 interaction_matrix
    (2000, 2000, 3, 3), FP64 and mixed, outputs left in HBM."""
    import mhs_amd as M
    from mhs_amd import api, workloads
    obs, tris, eps = workloads.c2_stress()
    mat = M.Material(30.0, 30.0)

    def stress(mode):
        with M.precision(mode):
            return api.plan(M.stress_matrix, obs, tris, mat, eps)

    def inter(mode):
        with M.precision(mode):
            return api.plan(M.interaction_matrix, tris, mat, eps)
    return {"stress": _modes(stress, obs.shape[0] * tris.shape[0], steps),
            "interaction_3x3": _modes(inter, tris.shape[0] * tris.shape[0], steps)}


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


#: Poisson arrival rates of the served stream, frozen: 0.7 / E[service] per
#: mode, E[service] from a per-(form, order) linear cost model (seconds per
#: request + seconds per pair-eval) fitted to the first 128 requests served
#: one at a time on an MI355X and applied to all 512 (mixed 0.0807 s, FP64
#: 0.175 s a request).
SERVE_RATE = {"mixed": 8.68, "fp64": 4.00}
#: The stated latency targets.  The stream's size tail sets p99: its largest
#: request (an order-2 3 x 3 interaction of a 7,482-triangle mesh, 146 GB of
#: output) takes 3.8 s mixed and 9.2 s FP64 alone, and its four requests of
#: 103-146 GB arrive within 10 s and cannot share the 230 GB arena, so they
#: run one after another.  A processor-sharing simulation of the stream on
#: the cost model with that memory limit and this admission policy gives p99
#: 5.5 s mixed and 14.5 s FP64; the targets leave 20 % above that.  p50 is
#: set by the many small requests: 0.25 s.
SERVE_P99_TARGET_S = {"mixed": 6.6, "fp64": 17.4}
SERVE_P50_TARGET_S = 0.25
SERVE_WARM = 48      # requests served one at a time first (tuning the classes they visit)


def served_stream(n_req: int, precision: str, queue: bool = False) -> dict:
    """The served stream at its frozen rate: the first SERVE_WARM requests
    served one at a time with tuning on (a warm-up, untimed), then ``n_req``
    requests with Poisson arrivals at SERVE_RATE through the continuous-
    batching loop (4 slots, uncut meshes admitted against the device budget,
    results dropped after completion); p50 / p99 latency, pair-evals/s and
    the p99 target's pass/fail.  ``queue``: the loop's queue mode (one
    stream, batches of every active request's next chunk, each pass drained
    from a device work queue by one persistent launch)."""
    from dataclasses import replace
    from mhs_amd import serve
    rate = SERVE_RATE[precision]
    reqs = serve.c5_requests(n_req=n_req, rate=rate, precision=precision)
    _progress(f"served stream {precision}: warm-up on {SERVE_WARM} requests")
    warm = serve.Server(slots=1, backend=serve.GpuBackend(1, tune=True))
    warm.run([replace(r, arrival=0.0) for r in reqs[:SERVE_WARM]], keep=False, realtime=False)
    sched = "queue" if queue else "streams"
    _progress(f"served stream {precision} ({sched}): {n_req} requests at {rate:.2f}/s")
    be = serve.GpuBackend(4, queue=queue)
    srv = serve.Server(slots=4, backend=be, queue=queue)
    out = srv.run(reqs, keep=False)
    summ = serve.summary(out)
    summ["loop_host_s"] = {k: round(v, 3) for k, v in srv.stats.items()}
    target = SERVE_P99_TARGET_S[precision]
    summ.update({"precision": precision, "scheduler": sched, "rate_per_s": rate, "rate": "frozen: 0.7 / E[service]",
                 "slots": 4, "batches": srv.batches if queue else None, "cache_trims": be.trims, "p99_target_s": target,
                 "p99_pass": bool(summ["p99_s"] <= target and summ["completed"] == n_req),
                 "p50_target_s": SERVE_P50_TARGET_S, "p50_pass": bool(summ["p50_s"] <= SERVE_P50_TARGET_S)})
    return summ


def _synthetic_tris(kind: str):
    """Branch-pure meshes for the FLOP-model check of the counter pass.
    "far": 4,096 triangles of 0.1 km on a 16^3 lattice 2 km apart, 5-35 km
    deep, each turned at random: every pair but the diagonal is a far pair
    (20-520 L apart) in both families.  "near": a 12 x 12-cell plane of
    1 km cells (288 triangles) 50 km deep, dip 40: most direct pairs within
    the 10 L switch, every image pair far.  Returns (tris, eps)."""
    from mhs_amd import workloads
    rng = np.random.default_rng(20261204)
    if kind == "far":
        g = np.stack(np.meshgrid(*(np.arange(16.0),) * 3, indexing="ij"), -1).reshape(-1, 3) * 2.0
        c = g + np.array([-15.0, -15.0, -35.0])
        shape = np.array([[0.06, 0.0, 0.0], [-0.03, 0.05, 0.0], [-0.03, -0.05, 0.0]])
        q = rng.normal(size=(c.shape[0], 3, 3))
        rot = np.linalg.qr(q)[0]
        tris = c[:, None, :] + np.einsum("nij,kj->nki", rot, shape)
        return workloads.orient_up(tris), 0.01
    if kind == "near":
        return workloads.orient_up(workloads.grid_mesh(workloads.plane(12.0, 12.0, 12, 12, 40.0, -50.0))), 0.1
    raise ValueError(kind)


def _job(mixed, workload: str = "c3"):
    from mhs_amd import runtime, tdcs, workloads
    from mhs_amd.api import collocation_points
    tris, eps = workloads.c3_interaction() if workload == "c3" else _synthetic_tris(workload)
    n = tris.shape[0]
    fr = tdcs.slip_frame(tris)
    return runtime.Job(kind="interaction", order=0, obs=collocation_points(tris, 0), tris=tris,
                       eps=np.full(n, eps), mu=30.0, lam=30.0, fold=True, subtract=True,
                       nrm=np.ascontiguousarray(fr[:, 2, :]), rb=np.ascontiguousarray(fr[:, 0:1, :]),
                       sb=np.ascontiguousarray(fr[:, 0:1, :]), mixed=mixed, self_rows=True)


def _time_mode(mixed, steps, warmup):
    """Steps of one mode: wall ms per step and per-pass HIP-event ms."""
    import torch
    from mhs_amd import _lib, runtime
    prep = runtime.Prepared(_job(mixed))
    st = prep.stream
    with torch.cuda.stream(st):
        for _ in range(warmup):
            prep.launch()
        torch.cuda.synchronize()
        prep.check_fold()
        t0 = time.perf_counter()
        for _ in range(steps):
            prep.launch()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        stage_ms = {}
        # the contracted interaction's passes: far (k_far), near (k_near), fold
        for name, mask in (("far", _lib.STAGE_Q), ("near", _lib.STAGE_PAIR), ("fold", _lib.STAGE_FOLD)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            prep.launch(stages=mask)
            e1.record(st)
            e1.synchronize()
            stage_ms[name] = e0.elapsed_time(e1)
        prep.launch()                    # leave a whole result for the error check
        torch.cuda.synchronize()
    stage_ms["far_tiles"] = prep.far_tiles
    stage_ms["tune_ms"] = {str(k): round(v, 3) for k, v in prep.dev.tune_times.get(prep.tune_key(), {}).items()}
    return prep, dt, stage_ms


def _served_all(args) -> dict:
    """The served stream in every requested precision and scheduler, keyed
    "<precision>" (a stream per request) or "<precision>_queue"."""
    out = {}
    for m in (x for x in args.serve_modes.split(",") if x):
        if args.serve_scheduler in ("streams", "both"):
            out[m] = served_stream(args.serve_requests, m)
        if args.serve_scheduler in ("queue", "both"):
            out[f"{m}_queue"] = served_stream(args.serve_requests, m, queue=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--cpu-cols", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=512,
                    help="requests of the served stream (0 = skip it)")
    ap.add_argument("--serve-modes", default="mixed,fp64",
                    help="precisions the served stream runs in, comma-separated")
    ap.add_argument("--serve-scheduler", default="streams", choices=["streams", "queue", "both"],
                    help="the serving loop's scheduler: a stream per request, the batched work "
                         "queue, or both on the same stream")
    ap.add_argument("--serve-only", action="store_true",
                    help="run the served stream alone and exit")
    ap.add_argument("--source-major", action="store_true",
                    help="time the device source-major copy of interaction-shaped matrices and exit")
    ap.add_argument("--workload", default="c3", choices=["c3", "far", "near"],
                    help="the counter pass's workload: the interaction n=10k, or a branch-pure "
                         "synthetic mesh (every pair far; most direct pairs near)")
    ap.add_argument("--c4", action="store_true",
                    help="time the three-fault network's continuous-layout order-2 interaction and exit")
    ap.add_argument("--c2", action="store_true",
                    help="time the Gaussian-field fault's stress and 3 x 3 interaction matrices and exit")
    ap.add_argument("--shared-nodes", action="store_true",
                    help="time the shared-node table, host against device, and exit")
    ap.add_argument("--counter-pass", action="store_true",
                    help="tune, run one whole step of --mode, then a fill and a copy of 256 MiB, and "
                         "print the step's modelled FLOPs per kernel (the program profiled by the "
                         "hardware counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the assembly shards by "
              f"receiver rows with no exchange, and no multi-GPU figure is produced here.",
              file=sys.stderr)
        return 2
    if args.serve_only:
        print(json.dumps({"served": _served_all(args)}), flush=True)
        return 0
    if args.counter_pass:
        counter_pass(args.mode, args.workload)
        return 0
    if args.shared_nodes:
        print(json.dumps({"shared_nodes": shared_nodes_timing()}))
        return 0
    if args.c4:
        print(json.dumps({"c4": c4_timing()}))
        return 0
    if args.c2:
        print(json.dumps({"c2": c2_timing()}))
        return 0
    if args.source_major:
        print(json.dumps({"source_major": source_major_timing()}))
        return 0

    from mhs_amd import flops, runtime
    mixed = args.mode == "mixed"
    _progress("timing the other mode")
    prep_o, dt_o, stage_o = _time_mode(not mixed, args.steps, args.warmup)
    ref = prep_o.final.clone()
    hist_o = prep_o.histogram()
    fold_o = prep_o.fold_histogram()
    del prep_o
    _progress(f"timing {args.mode}")
    prep, dt, stage_ms = _time_mode(mixed, args.steps, args.warmup)
    hist = prep.histogram()
    fold_h = prep.fold_histogram()
    a, b = (prep.final, ref) if mixed else (ref, prep.final)
    d = (a - b).abs()
    relerr = float((d.max() / b.abs().max()).item())
    # per receiver row: the row's largest error over the row's largest entry
    row_relerr = float((d.amax(dim=(1, 2, 3)) / b.abs().amax(dim=(1, 2, 3))).max().item())
    rule_key = prep.rule_key()          # the mixed run's entry (either run's key)
    del prep, ref, a, b, d

    n = 10_000
    pairs = n * n
    value = pairs / dt
    h_mixed = hist if mixed else hist_o
    h_fp64 = hist_o if mixed else hist
    fl = flops.flops(h_fp64)
    flm = flops.flops_mixed(h_mixed)

    def roofline(kernel, f32, f64, ms):
        achieved = (f32 + f64) / (ms * 1e-3) / 1e12
        peak = (f32 + f64) / (f32 / PEAK_TFLOPS["f32"] + f64 / PEAK_TFLOPS["f64"])
        return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "flop": f32 + f64, "flop_fp32": f32, "ms": ms}

    st_m, st_d = (stage_ms, stage_o) if mixed else (stage_o, stage_ms)
    roof_mixed = roofline("k_far_mixed (direct + image R + Q far field at the mixed rule, FP32 points)",
                          flm["far_fp32"], flm["far_fp64"], st_m["far"])
    roof_fp64 = roofline("k_far (direct + image R + Q far field at the FP64 rule)", 0.0, fl["far"], st_d["far"])
    near_roofs = [roofline("k_near (closed-form near field, mixed rule's switch)", 0.0, flm["near"], st_m["near"]),
                  roofline("k_near (closed-form near field, FP64 rule's switch)", 0.0, fl["near"], st_d["near"])]
    fh_m, fh_d = (fold_h, fold_o) if mixed else (fold_o, fold_h)
    ff_m, ff_d = flops.flops_fold(fh_m, True), flops.flops_fold(fh_d, False)
    fold_roofs = [roofline("k_fold (near and far points; list and grids included), mixed rule",
                           ff_m["fp32"], ff_m["fp64"], st_m["fold"]),
                  roofline("k_fold (near and far points; list and grids included), FP64 rule",
                           0.0, ff_d["fp64"], st_d["fold"])]
    main_roof = roof_mixed if mixed else roof_fp64
    dev = runtime.device()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "interaction_matrix receiver=source='strike', n=10,000 triangles, vertical "
                               "strike-slip fault 100 x 15 km reaching z=0 (fold active), seeded 5% vertex "
                               "jitter, eps=0.1, Material(30, 30), order 0",
                   "n": n, "pairs": pairs, "mode": args.mode},
        "roofline": dict(main_roof, traffic=_traffic(args.mode)),
        "roofline_other_kernels": [roof_fp64 if mixed else roof_mixed] + near_roofs + fold_roofs,
        "stage_ms": stage_ms,
        "other_mode": {"mode": "fp64" if mixed else "mixed", "ms_per_step": dt_o * 1e3,
                       "value": pairs / dt_o, "stage_ms": stage_o},
        "mixed_over_fp64": (dt_o / dt) if mixed else (dt / dt_o),
        "mixed_relmax_vs_fp64": relerr,
        "mixed_row_relmax_vs_fp64": row_relerr,
        "mixed_rule": {"chosen": runtime._rule_name(dev.rules[rule_key]) if rule_key in dev.rules else "default",
                       "tuner": dev.rule_table.get(rule_key, {}),
                       "bound": runtime.RULE_ERROR_BOUND},
        "histogram": hist if mixed else hist_o,
        "flop": {"fp64_rule": fl, "mixed_rule": flm, "fold_mixed_rule": ff_m, "fold_fp64_rule": ff_d},
        "fold_histogram": {"mixed_rule": fh_m, "fp64_rule": fh_d},
    }
    if args.serve_requests > 0:
        from mhs_amd import serve
        line["served"] = _served_all(args)
        serve.release_arenas()
    if not args.no_cpu:
        _progress(f"CPU baseline on {CPU_SHARE} processes")
        line["cpu_baseline"] = cpu_baseline(args.cpu_cols)
        line["vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
