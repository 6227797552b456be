"""The per-pair device math built for the host under AddressSanitizer and
UndefinedBehaviorSanitizer, every function inlined into its caller (the
variant whose device build once gave lane-dependent NaNs): the direct nodal
stress, blob weight and displacement (edge integrals in all their regimes,
closed-form near moments, Gauss far tables, the lift) and the image
R-family, per (observer, source, node), against the oracle
(fullspace.nodal + image.image_r: src/mhs/fullspace/kernels.py:126-213,
src/mhs/kernels/image.py:435-497).  Any out-of-bounds access or undefined
operation aborts the harness (no recovery)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"
HARNESS = ROOT / "tests" / "host_core" / "core_harness.hip"
MU, LAM = 30.0, 45.0


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = tmp_path_factory.mktemp("core") / "core_inlined"
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O0", "-g", "-std=c++17",
           "-Xarch_host", "-fsanitize=address,undefined", "-Xarch_host", "-fno-sanitize-recover=all",
           "-DMHS_EDGE_FN=MHS_INLINE", "-DMHS_DIRECT_FN=MHS_INLINE", "-DMHS_NEAR_FN=MHS_INLINE",
           str(HARNESS), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def _sources():
    """Buried sources (no fold) of several shapes and orientations."""
    rng = np.random.default_rng(11)
    tris = [np.array([[0.0, 0.0, -2.0], [1.0, 0.0, -2.0], [0.2, 0.9, -2.0]]),          # horizontal
            np.array([[0.0, 0.0, -1.5], [1.2, 0.1, -1.5], [0.6, 0.0, -2.6]]),           # steep
            np.array([[0.0, 0.0, -3.0], [2.0, 0.0, -3.0], [1.9, 0.15, -3.4]])]          # thin
    for _ in range(2):
        c = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), -rng.uniform(1.2, 3.0)])
        tris.append(c + rng.normal(scale=0.5, size=(3, 3)) * [1, 1, 0.3])
    return [t for t in tris if t[:, 2].max() < -0.5]


def _observers(tri, rng):
    """Observers in every regime of one source: on and just off its plane
    over the interior, at vertices and edge midpoints, above the centroid,
    beyond the 10 L switch (12 x 12 and 8 x 8 tables), and near the free
    surface over the mirrored triangle."""
    c = tri.mean(axis=0)
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    n /= np.linalg.norm(n)
    L = np.sqrt(0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0])))
    pts = [c, c + 1e-3 * n, c + 0.05 * n, c + 0.5 * n, c - 2.0 * n]
    for v in range(3):
        pts += [tri[v], tri[v] + 0.01 * n, 0.5 * (tri[v] + tri[(v + 1) % 3]) + 0.02 * n]
        pts.append(tri[v] + 0.3 * (tri[v] - c))
    for d in (3.0, 9.0, 15.0, 60.0):
        u = rng.normal(size=3)
        pts.append(c + d * L * u / np.linalg.norm(u))
    pts += [[c[0], c[1], 0.0], [c[0] + 0.3, c[1] - 0.2, -0.05]]
    P = np.array(pts, float)
    P[:, 2] = np.minimum(P[:, 2], 0.0)
    return P


def _run(exe, P, cases):
    from mhs_amd import geometry
    gx, gw = geometry.gauss_tables()
    blob = [np.array([P, len(cases)], np.int32).tobytes(), np.array([MU, LAM]).tobytes(),
            gx.tobytes(), gw.tobytes()]
    for rec, o, k in cases:
        blob += [np.ascontiguousarray(rec, np.float64).tobytes(), np.asarray(o, np.float64).tobytes(),
                 np.array([k], np.int32).tobytes()]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe)], input=b"".join(blob), capture_output=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-4000:]
    out = np.frombuffer(r.stdout, np.float64).reshape(len(cases), 37)
    return out[:, :27].reshape(-1, 3, 3, 3), out[:, 27], out[:, 28:].reshape(-1, 3, 3)


def _image_quadrature(obs, tri, P, eps, monkeypatch):
    """The image R-family by the oracle's Gauss tables at 200 x 200 points
    (converged to ~1e-15 here: the mirrored triangle lies above z = 0, the
    observers below it)."""
    from oracle import image, moments
    monkeypatch.setattr(moments, "FAR_GAUSS_N", 200)
    monkeypatch.setattr(moments, "FAR_GAUSS_N_DISTANT", 200)
    try:
        return image.image_r(obs, tri, P, MU, LAM, eps, ("U", "H"), far_field="quadrature")
    finally:
        monkeypatch.undo()


@pytest.mark.parametrize("P", [0, 1, 2])
def test_pair_math_sanitized_and_against_oracle(harness, P, monkeypatch):
    """Parity bound per source column block: 1e-12 against the oracle.  In
    one band the reference's own closed form is ill-conditioned: the image
    R-family of a mirrored triangle 2.5-10 L away at P >= 1 (its raw
    moments to degree 7 carry relative errors up to 1e-4 that cancel in the
    contraction), where both restatements stray from the converged integral
    by up to ~1e-10 of the block (measured: this build 7.6e-11, the oracle
    1.8e-11, on one P2 observer); there the expected value is the oracle's
    direct term plus the image R-family by converged quadrature, at 2e-10."""
    from mhs_amd import geometry
    from oracle import fullspace, image
    rng = np.random.default_rng(100 + P)
    K = (P + 1) * (P + 2) // 2
    nu = 0.5 * LAM / (LAM + MU)
    eps = 0.1
    worst = {"H": 0.0, "E": 0.0, "U": 0.0}
    worst_band = dict(worst)
    for tri in _sources():
        obs = _observers(tri, rng)
        rec = geometry.source_records(tri[None], np.array([eps]))[0]
        cases = [(rec, o, k) for o in obs for k in range(K)]
        H, E, U = _run(harness, P, cases)
        d = fullspace.nodal(obs, tri, P, MU, nu, eps, want=("U", "H", "E"))
        im = image.image_r(obs, tri, P, MU, LAM, eps, ("U", "H"))
        D_img = np.sqrt(np.sum((obs - rec[geometry.S_IC:geometry.S_IC + 3]) ** 2, axis=1) + eps * eps)
        band = (D_img > 2.5 * rec[geometry.S_IL]) & (D_img <= 10.0 * rec[geometry.S_IL]) & (P >= 1)
        if band.any():
            imq = _image_quadrature(obs[band], tri, P, eps, monkeypatch)
            for key in ("U", "H"):
                im[key][band] = imq[key]
        want = {"H": (d["H"] + im["H"]).reshape(len(obs), -1), "E": d["E"].reshape(len(obs), -1),
                "U": (d["U"] + im["U"]).reshape(len(obs), -1)}
        got = {"H": H.reshape(len(obs), -1), "E": E.reshape(len(obs), -1), "U": U.reshape(len(obs), -1)}
        for key in worst:
            g, w = got[key], want[key]
            assert np.all(np.isfinite(g)), key
            err = np.abs(g - w).max(axis=1) / np.abs(w).max()
            worst[key] = max(worst[key], float(err[~band].max()))
            if band.any():
                worst_band[key] = max(worst_band[key], float(err[band].max()))
    print(f"P{P} relmax per source block vs oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items())
          + "; image 2.5-10 L band vs converged quadrature: "
          + ", ".join(f"{k} {v:.2e}" for k, v in worst_band.items()))
    assert max(worst.values()) < 1e-12, worst
    assert max(worst_band.values()) < 2e-10, worst_band
