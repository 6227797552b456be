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
of the reference) on every receiver against a seeded sample of source
columns, the columns split over this host's cores.  Prints ONE JSON line.
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


def _cpu_block(cols):
    from mhs_amd import workloads
    from oracle import api as oapi
    tris, eps = workloads.c3_interaction()
    K = oapi.interaction_matrix(tris[cols], 30.0, 30.0, eps, receiver="strike", source="strike",
                                obs_tris=tris)
    return int(K.shape[0] * K.shape[1])


def cpu_baseline(n_cols: int, cores: int) -> dict:
    """The FP64 oracle on all 10^4 receivers x ``n_cols`` seeded source
    columns (the reference's own loop runs per source, vectorised over the
    receivers), one column per task on a pool of one spawned process per core
    (BLAS pinned to one thread), the surface-reaching (folded, costliest)
    columns first; pair-evals/s over the wall time.  The sample is
    stratified: surface-reaching sources (shallowest vertex above 3 eps
    depth) and the rest are drawn in proportion to their counts."""
    import multiprocessing as mp
    from mhs_amd import workloads
    rng = np.random.default_rng(CPU_SEED)
    tris, eps = workloads.c3_interaction()
    shallow = tris[:, :, 2].max(axis=1) > -3.0 * eps
    strad, rest = np.flatnonzero(shallow), np.flatnonzero(~shallow)
    n_sh = int(round(n_cols * len(strad) / len(tris)))
    cols = np.concatenate([rng.choice(strad, size=n_sh, replace=False),
                           rng.choice(rest, size=n_cols - n_sh, replace=False)])
    blocks = [np.array([c]) for c in cols]
    env = {"OMP_NUM_THREADS": "1", "OPENBLAS_NUM_THREADS": "1", "MKL_NUM_THREADS": "1"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = mp.get_context("spawn")
        with ctx.Pool(min(cores, len(blocks))) as pool:
            t0 = time.perf_counter()
            pairs = sum(pool.imap_unordered(_cpu_block, blocks, chunksize=1))
            dt = time.perf_counter() - t0
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    used = min(cores, len(blocks))
    return {"value": pairs / dt, "unit": "pair-evals/s", "cores": used, "kind": "port",
            "sample": f"all 10,000 receivers x {n_cols} seeded source columns (seed {CPU_SEED}; {n_sh} of them "
                      f"surface-reaching, as in the whole mesh) of the same workload, FP64 oracle, "
                      f"{pairs:.3e} pairs (their fold pairs included) in "
                      f"{dt:.1f} s wall on {used} processes"}


def _traffic(mode: str):
    """HBM bytes per far-pass launch from the committed counter runs
    (FETCH_SIZE, WRITE_SIZE, each scaled by a copy / fill of known size), or
    None when they were taken in the other mode or are absent."""
    path = os.path.join(ROOT, "profiles", "k_far_mixed_traffic.json")
    if mode != "mixed" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["hbm_bytes_per_launch"]


def counter_pass(mode: str) -> None:
    """One far pass of the workload (autotuner off, default column tiles),
    then a 256 MiB fill and a 256 MiB device copy: the program that hardware
    counter runs (FETCH_SIZE, WRITE_SIZE) profile, the fill and copy being
    the known-size calibration of those counters."""
    import torch
    from mhs_amd import _lib, runtime
    runtime.autotune = False
    prep = runtime.Prepared(_job(mode == "mixed"))
    with torch.cuda.stream(prep.stream):
        prep.launch(stages=_lib.STAGE_Q)
    torch.cuda.synchronize()
    a = torch.empty(1 << 25, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    _progress(f"counter pass done ({mode}, default column tiles per block)")


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


def served_stream(n_req: int, precision: str, n_cal: int = 48) -> dict:
    """The served stream: calibrate E[service] on the first ``n_cal`` requests
    run one at a time, set the Poisson rate to 0.7 / E[service], then serve
    ``n_req`` requests through the continuous-batching loop (4 slots, results
    dropped after completion); p50 / p99 latency and pair-evals/s."""
    from dataclasses import replace
    from mhs_amd import serve
    _progress(f"served stream: calibrating on {n_cal} requests")
    cal = serve.c5_requests(n_req=n_cal, rate=1.0, precision=precision)
    res = serve.Server(slots=1).run([replace(r, arrival=0.0) for r in cal], keep=False, realtime=False)
    service = float(np.mean([s.done - s.admitted for s in res if s.error is None]))
    rate = 0.7 / service
    _progress(f"served stream: {n_req} requests at {rate:.2f}/s")
    reqs = serve.c5_requests(n_req=n_req, rate=rate, precision=precision)
    out = serve.Server(slots=4).run(reqs, keep=False)
    summ = serve.summary(out)
    summ.update({"precision": precision, "rate_per_s": rate, "mean_service_s": service,
                 "calibration": f"first {n_cal} requests of the same seeded stream, one at a time",
                 "slots": 4})
    return summ


def _job(mixed):
    from mhs_amd import runtime, tdcs, workloads
    from mhs_amd.api import collocation_points
    tris, eps = workloads.c3_interaction()
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
    ap.add_argument("--source-major", action="store_true",
                    help="time the device source-major copy of interaction-shaped matrices and exit")
    ap.add_argument("--c4", action="store_true",
                    help="time the three-fault network's continuous-layout order-2 interaction and exit")
    ap.add_argument("--c2", action="store_true",
                    help="time the Gaussian-field fault's stress and 3 x 3 interaction matrices and exit")
    ap.add_argument("--shared-nodes", action="store_true",
                    help="time the shared-node table, host against device, and exit")
    ap.add_argument("--counter-pass", action="store_true",
                    help="run one far pass of --mode plus a fill and a copy of 256 MiB, print nothing "
                         "(the program profiled by the FETCH_SIZE / WRITE_SIZE counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the assembly shards by "
              f"receiver rows with no exchange, and no multi-GPU figure is produced here.",
              file=sys.stderr)
        return 2
    if args.counter_pass:
        counter_pass(args.mode)
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
    del prep_o
    _progress(f"timing {args.mode}")
    prep, dt, stage_ms = _time_mode(mixed, args.steps, args.warmup)
    hist = prep.histogram()
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
        "roofline_other_kernels": [roof_fp64 if mixed else roof_mixed] + near_roofs,
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
        "flop": {"fp64_rule": fl, "mixed_rule": flm},
    }
    if args.serve_requests > 0:
        line["served"] = served_stream(args.serve_requests, args.mode)
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, 16)
        _progress(f"CPU baseline on {cores} processes")
        line["cpu_baseline"] = cpu_baseline(args.cpu_cols, cores)
        line["vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
