"""Write the golden vectors of the reference implementation (build container only).

Imports brendanjmeade/mhs from /root/reference/src -- which exists only in the
build container -- and records inputs and outputs as small .npz fixtures next
to this script.  Only functions whose value no mathematical check pins to
1e-12 are recorded here: the end-to-end matrices of every entry point, order
and layout (they fix the near/far branch choices), the Q-family orders, the
fold's polar grid and folded image.  The fixtures are data; the reference's
code does not travel.

Run:  python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, "/root/reference/src")

import mhs  # noqa: E402  (the reference, build container only)
from mhs.kernels import fold as ref_fold  # noqa: E402
from mhs.kernels import image as ref_image  # noqa: E402

from mhs_amd import workloads  # noqa: E402


def tiny_buried():
    """8 triangles of a 30-deg dipping, strike-20-deg plane, top at -1.5 km,
    perturbed; per-source eps."""
    rng = np.random.default_rng(11)
    P = workloads.plane(1.6, 1.6, 2, 2, 30.0, -1.5, 0.2, -0.1, 20.0)
    P = P + rng.uniform(-0.05, 0.05, P.shape)
    tris = workloads.orient_up(workloads.grid_mesh(P))
    eps = rng.uniform(0.08, 0.15, tris.shape[0])
    return tris, eps


def tiny_surface():
    """12 triangles of a vertical fault x in [-0.75, 0.75], z in [-1.2, 0]:
    the top row touches z = 0 and is folded."""
    x = np.linspace(-0.75, 0.75, 4)
    z = np.linspace(0.0, -1.2, 3)
    X, Z = np.meshgrid(x, z, indexing="ij")
    P = np.stack([X, 0.02 * X * Z, Z], axis=-1)
    return workloads.grid_mesh(P), 0.1


def obs_set(rng, tris, n_near=10, n_mid=8, n_far=6, n_surf=6):
    c = tris.reshape(-1, 3).mean(axis=0)
    span = np.ptp(tris.reshape(-1, 3), axis=0).max()
    pts = []
    cen = tris.mean(axis=1)
    for _ in range(n_near):
        s = rng.integers(tris.shape[0])
        pts.append(cen[s] + rng.normal(0, 0.15, 3))
    for _ in range(n_mid):
        pts.append(c + rng.normal(0, 2.0 * span, 3))
    for _ in range(n_far):
        v = rng.normal(0, 1, 3)
        pts.append(c + v / np.linalg.norm(v) * rng.uniform(15, 60) * span)
    for _ in range(n_surf):
        pts.append(np.array([c[0] + rng.normal(0, 1.5), c[1] + rng.normal(0, 1.5), 0.0]))
    P = np.array(pts)
    P[:, 2] = np.minimum(P[:, 2], 0.0)
    return P


def band_obs(rng, n=10):
    """Receivers in and around the fold band of tiny_surface."""
    pts = np.column_stack([rng.uniform(-2.0, 2.0, n), rng.uniform(-0.6, 0.6, n),
                           -rng.uniform(0.0, 1.4, n)])
    pts[0] = [0.1, 0.05, 0.0]
    pts[1] = [0.3, -0.02, -0.05]
    return pts


def save(name, **arrays):
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **arrays)
    print(f"  wrote {path.name} ({path.stat().st_size / 1024:.0f} KiB)", flush=True)


def case_small():
    rng = np.random.default_rng(5)
    mat = mhs.Material(30.0, 45.0)
    tb, eb = tiny_buried()
    ts, es = tiny_surface()
    ob = obs_set(rng, tb)
    os_ = np.concatenate([band_obs(rng), obs_set(rng, ts, 4, 4, 3, 3)])
    out = {"tb": tb, "eb": eb, "ts": ts, "es": np.array(es), "ob": ob, "os": os_,
           "mu": np.array(mat.mu), "lam": np.array(mat.lam)}
    nrm_b = rng.normal(size=(ob.shape[0], 3))
    nrm_s = rng.normal(size=(os_.shape[0], 3))
    out["nrm_b"], out["nrm_s"] = nrm_b, nrm_s
    t0 = time.time()
    for p in (0, 1, 2):
        out[f"disp_b_p{p}"] = mhs.disp_matrix(ob, tb, mat, eb, order=p)
        out[f"stress_b_p{p}"], out[f"eig_b_p{p}"] = mhs.stress_matrix(ob, tb, mat, eb, order=p, parts=True)
        out[f"trac_b_p{p}"] = mhs.traction_matrix(ob, nrm_b, tb, mat, eb, order=p)
        out[f"inter_b_p{p}"] = mhs.interaction_matrix(tb, mat, eb, order=p)
    print(f"  buried done {time.time() - t0:.1f}s", flush=True)
    out["total_b_p0"] = mhs.total_stress_matrix(ob, tb, mat, eb)
    out["strain_b_p0"] = mhs.elastic_strain_matrix(ob, tb, mat, eb)
    out["trac_b_noeig"] = mhs.traction_matrix(ob, nrm_b, tb, mat, eb, subtract_eigenstress=False)
    out["inter_b_ss"] = mhs.interaction_matrix(tb, mat, eb, receiver="strike", source="strike")
    out["inter_b_rows"] = mhs.interaction_matrix(tb, mat, eb, order=1, rows=(5, 17),
                                                 receiver=("dip", "normal"), source="tensile")
    out["inter_b_obs"] = mhs.interaction_matrix(tb, mat, eb, obs_tris=tb[:3], shrink=0.2, order=1)
    # continuous layout
    sn = mhs.shared_nodes(tb)
    out["sn_coords"], out["sn_ids"], out["sn_normals"] = sn.coords, sn.ids, sn.normals
    out["sn_weights"], out["sn_boundary"] = sn.weights, sn.boundary
    out["inter_b_cont"] = mhs.interaction_matrix(tb, mat, eb, order=2, layout="continuous")
    out["disp_b_cont"] = mhs.disp_matrix(ob, tb, mat, eb, order=2, layout="continuous")
    out["trac_b_cont"] = mhs.traction_matrix(ob, nrm_b, tb, mat, eb, order=2, layout="continuous")
    out["frame_b"] = mhs.tdcs.slip_frame(tb)
    print(f"  buried extras done {time.time() - t0:.1f}s", flush=True)
    # surface-reaching (fold)
    for p in (0, 2):
        out[f"disp_s_p{p}"] = mhs.disp_matrix(os_, ts, mat, es, order=p)
        out[f"trac_s_p{p}"] = mhs.traction_matrix(os_, nrm_s, ts, mat, es, order=p)
        out[f"inter_s_p{p}"] = mhs.interaction_matrix(ts, mat, es, order=p)
        print(f"  surface p{p} done {time.time() - t0:.1f}s", flush=True)
    out["stress_s_p0"], out["eig_s_p0"] = mhs.stress_matrix(os_, ts, mat, es, parts=True)
    out["disp_s_nofold"] = mhs.disp_matrix(os_, ts, mat, es, fold=False)
    out["stress_s_nofold"] = mhs.stress_matrix(os_, ts, mat, es, fold=False)
    out["inter_s_cont"] = mhs.interaction_matrix(ts, mat, es, order=2, layout="continuous",
                                                 receiver="strike", source=("strike", "dip"))
    out["frame_s"] = mhs.tdcs.slip_frame(ts)
    print(f"  surface done {time.time() - t0:.1f}s", flush=True)
    save("small", **out)


def case_qorders():
    rng = np.random.default_rng(7)
    tris, _ = tiny_surface()
    obs = np.concatenate([band_obs(rng, 40), rng.uniform(-3, 3, (40, 3)) * [1, 1, 0.5] - [0, 0, 1.5]])
    obs[:, 2] = np.minimum(obs[:, 2], 0.0)
    orders = []
    for s in range(tris.shape[0]):
        fr = mhs.fullspace.frame.local_frame(tris[s])
        for p in (0, 2):
            for eps in (0.01, 0.1):
                orders.append(ref_image.q_gauss_orders(obs, tris[s] * np.array([1, 1, -1.0]),
                                                       eps, fr.L, p))
    save("qorders", obs=obs, tris=tris, orders=np.array(orders))


def case_fold():
    rng = np.random.default_rng(9)
    mat = mhs.Material(30.0, 45.0)
    tris, eps = tiny_surface()
    obs = band_obs(rng, 6)
    s_idx = [0, 1, 6]
    out = {"tris": tris, "obs": obs, "s_idx": np.array(s_idx)}
    for s in s_idx:
        c, r = ref_fold.crease(tris[s], eps)
        out[f"crease_{s}"] = np.array([c[0], c[1], r])
        diam = max(float(np.linalg.norm(tris[s][i] - tris[s][j])) for i in range(3) for j in range(3))
        for o in range(2):
            pts, wts, R1 = ref_fold.polar_grid(obs[o, :2], max(0.0, -obs[o, 2]), c, r, eps, diam)
            out[f"grid_{s}_{o}_pts"], out[f"grid_{s}_{o}_wts"] = pts, wts
            out[f"grid_{s}_{o}_R1"] = np.array(R1)
        for p in (0, 1):
            img = ref_fold.image(obs, tris[s], mat.mu, mat.lam, eps, p, ("U", "H"))
            out[f"img_{s}_p{p}_U"], out[f"img_{s}_p{p}_H"] = img["U"], img["H"]
    save("fold", **out)


def case_c1():
    obs, tris, eps = workloads.c1_disp()
    mat = mhs.Material(30.0, 30.0)
    sel = np.arange(0, obs.shape[0], 16)
    G = mhs.disp_matrix(obs[sel], tris, mat, eps)
    save("c1_rows", sel=sel, G=G)


def case_c3():
    tris, eps = workloads.c3_interaction()
    mat = mhs.Material(30.0, 30.0)
    out = {}
    for lo in (0, 5000, 5024, 9998):
        t0 = time.time()
        out[f"rows_{lo}"] = mhs.interaction_matrix(tris, mat, eps, receiver="strike",
                                                   source="strike", rows=(lo, lo + 2))
        print(f"  c3 rows {lo} {time.time() - t0:.1f}s", flush=True)
    save("c3_rows", **out)


def case_c2():
    """Rows of the 2,000-triangle Gaussian-field fault: stress_matrix at a
    few observers (deep, mid-depth and two shallow ones inside the fold band
    of the top row) and row blocks of the default 3 x 3 interaction_matrix
    (an edge top-row pair, a mid-fault top-row pair in the band, a mid-depth
    pair and the last pair)."""
    obs, tris, eps = workloads.c2_stress()
    mat = mhs.Material(30.0, 30.0)
    shallow = np.flatnonzero((obs[:, 2] > -1.2) & (np.abs(obs[:, 1]) < 3.0) & (np.abs(obs[:, 0]) < 15.0))
    sel = np.concatenate([np.arange(4), shallow[:2]])
    t0 = time.time()
    # the inputs travel with the rows: the Gaussian field comes from a
    # Cholesky factor whose last bits depend on the host's BLAS
    out = {"obs": obs, "tris": tris, "eps": eps, "sel": sel,
           "stress": mhs.stress_matrix(obs[sel], tris, mat, eps)}
    print(f"  c2 stress rows {time.time() - t0:.1f}s", flush=True)
    for lo in (0, 1000, 1010, 1998):
        t0 = time.time()
        out[f"inter_{lo}"] = mhs.interaction_matrix(tris, mat, eps, rows=(lo, lo + 2))
        print(f"  c2 interaction rows {lo} {time.time() - t0:.1f}s", flush=True)
    save("c2_rows", **out)


def case_c4():
    """Row blocks of the three-fault continuous-layout order-2 interaction at
    the default receiver (strike, dip, normal) and source (strike, dip):
    (rows, V + E, 3, 2)."""
    tris, groups, eps = workloads.c4_network()
    mat = mhs.Material(30.0, 30.0)
    sn = mhs.shared_nodes(tris, groups=groups)
    out = {"tris": tris, "groups": groups, "eps": eps, "n_nodes": np.array(sn.n)}
    for lo in (0, 6000, sn.n - 2):
        t0 = time.time()
        out[f"rows_{lo}"] = mhs.interaction_matrix(tris, mat, eps, order=2, layout="continuous", shared=sn,
                                                   source=("strike", "dip"), rows=(lo, lo + 2))
        out["last_lo"] = np.array(lo)
        print(f"  c4 rows {lo} {time.time() - t0:.1f}s", flush=True)
    save("c4_rows", **out)


def case_table():
    """Fingerprint of the reference's generated image table, read as data:
    sha256 of the canonical JSON of its four record lists."""
    import ast
    import hashlib
    import json
    src = pathlib.Path("/root/reference/src/mhs/kernels/_image_table.py").read_text()
    vals = {n.targets[0].id: ast.literal_eval(n.value)
            for n in ast.parse(src).body if isinstance(n, ast.Assign)}
    out = {}
    for ref, mine in (("DG_RECORDS", "r_dg"), ("DDG_RECORDS", "r_ddg"),
                      ("Q_DG_RECORDS", "q_dg"), ("Q_DDG_RECORDS", "q_ddg")):
        recs = [[*r[:-1], [list(t) for t in r[-1]]] for r in vals[ref]]
        text = json.dumps(recs, separators=(",", ":"))
        out[mine] = {"n": len(recs), "sha256": hashlib.sha256(text.encode()).hexdigest()}
    (HERE / "table_fingerprint.json").write_text(json.dumps(out, indent=1) + "\n")
    print("  wrote table_fingerprint.json", flush=True)


CASES = {"table": case_table, "small": case_small, "qorders": case_qorders, "fold": case_fold,
         "c1": case_c1, "c2": case_c2, "c3": case_c3, "c4": case_c4}

if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for nm in names:
        print(f"case {nm}", flush=True)
        CASES[nm]()
