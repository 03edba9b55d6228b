"""Writes the golden fixtures in tests/golden/ from the compiled reference.

Run in a container that has the reference tree (it builds oracle/_ref with
`make -C oracle ref` first):

    python tests/golden/make_golden.py [--skip-large]

Inputs are synthetic configurations from gfc.generate_smooth (seeded, so the
GPU tests regenerate them bit for bit and check the stored SHA-256).  The
reference itself is run through oracle/ref_golden.c, a harness that calls
its exported functions; every number stored here was produced by the
reference's code.  Outputs:

  prims.npz        primitive-level vectors (det, inverse, projectSU3,
                   power3x3Binomial, updateSubLosAlamos, decomposeAlgebraSU3,
                   prod3x3) on 256 random and near-SU(3) matrices
  lat_4x8.npz      4^3x8, eps 0.25, seed 1: observables before/after the
                   initial reunitarization, the reunitarized U, e2/theta/F/
                   plaquettes after k = 1..200 half-sweeps, G after k in
                   KEEP_G, U after 200, runs to tolerance (estimate 1000 and 0)
  lat_8x16.npz     8^3x16, eps 0.25, seed 2: observables after
                   k = 1..200, G (sampled sites) after 200, runs to tolerance
  lat_2x2.npz      2^3x2 (eps 0.3, seed 11) and 6^3x4 (eps 0.3, seed 12),
  lat_6x4.npz      the same records as lat_4x8.npz
  lat_32x64.npz    32^3x64, eps 0.3, seed 3: observables after
                   every half-sweep k = 1..200 (harness mode `trace`, checked
                   bit for bit against `traj` on the small lattices), G
                   (sampled sites) after 200, run to tolerance
  lat_degen_4x8.npz  4^3x8 smooth configuration whose spatial links on
                   timeslices 0-2 are exact patterns that drive the local
                   update into the reference's degenerate branches (see
                   degenerate_config); trajectories for omega 1.5 and 1.0:
                   observables after every half-sweep k = 1..200, G and U
                   after k in KEEP_G
  io.npz           loadConfig results (SHA-256 of the loaded arrays and return
                   codes) for big-endian FP64, little-endian FP64 and
                   little-endian FP32 files, and for short / long files
  params.json      initORGaugeFixingParameters on sample parameter files
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "gaugefix-coulomb_amd"))
import gfc  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref")
KEEP_G = [1, 2, 3, 10, 11, 200]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(binary: str, *args) -> np.ndarray:
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        out = f.name
    cmd = [os.path.join(REF, binary), *map(str, args), out]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    data = np.fromfile(out, dtype="<f8")
    os.unlink(out)
    return data


def cfg_file(tmp: str, geom: gfc.Geometry, U: np.ndarray, prec: int, swap: bool) -> str:
    path = os.path.join(tmp, f"cfg_{geom.ns}x{geom.nt}_{prec}_{int(swap)}.bin")
    assert gfc.write_config(path, geom, U, prec, swap) == 0
    return path


def sampled_sites(V: int) -> np.ndarray:
    return np.arange(0, V, 7)[:4096] if V > 4096 else np.arange(V)


def parse_traj(data: np.ndarray, V: int, ks):
    full = V <= 4096
    nG = V * 18 if full else len(sampled_sites(V)) * 18
    nU = V * 4 * 18 if full else 0
    rec = 2 + 5 + nG + nU
    assert data.size == rec * len(ks), (data.size, rec, len(ks))
    out = []
    for i, _ in enumerate(ks):
        r = data[i * rec:(i + 1) * rec]
        obs = r[2:7]
        G = r[7:7 + nG].view(np.complex128).reshape(-1, 3, 3)
        U = r[7 + nG:].view(np.complex128).reshape(-1, 4, 3, 3) if nU else None
        out.append((int(r[0]), int(r[1]), obs, G, U))
    return out


def parse_full(data: np.ndarray, V: int):
    rc = int(data[0])
    obs = data[1:6]
    G = data[6:].view(np.complex128).reshape(-1, 3, 3)
    return rc, obs, G


def trace(harness: str, path: str, ns: int, nt: int, kmax: int) -> np.ndarray:
    """(k, e2, theta, F, plaq spatial, plaq temporal) after k = 1..kmax."""
    d = run(harness, "trace", path, ns, nt, 10, kmax).reshape(-1, 6)
    assert list(d[:, 0].astype(int)) == list(range(1, kmax + 1))
    return d


def lattice(tmp: str, name: str, ns: int, nt: int, eps: float, seed: int, ks, keep_g,
            harness: str, full_runs=((1000, 1e-16), (0, 1e-16)), keep_u=(200,),
            trace_kmax: int = 0):
    geom = gfc.Geometry(ns, nt)
    V = geom.volume
    U = gfc.generate_smooth(geom, seed, eps)
    path = cfg_file(tmp, geom, U, 8, False)
    fx = {"ns": ns, "nt": nt, "eps": eps, "seed": seed, "config_sha256": sha(U),
          "R": 10, "omega": np.float32(1.95), "hits": 2}
    o = run(harness, "obs", path, ns, nt)
    fx["obs_raw"] = o[0:5]
    fx["reunit_status"] = o[5]
    fx["obs_reunit"] = o[6:11]
    Ur = o[11:].view(np.complex128).reshape(V, 4, 3, 3)
    fx["U_reunit_sha256"] = sha(Ur)
    if V <= 512:
        fx["U_reunit"] = Ur
    ks = list(ks)
    t = parse_traj(run(harness, "traj", path, ns, nt, 10, *ks), V, ks)
    fx["traj_k"] = np.array([r[0] for r in t])
    fx["traj_rc"] = np.array([r[1] for r in t])
    fx["traj_obs"] = np.array([r[2] for r in t])
    for k, _, _, G, Uk in t:
        if k in keep_g:
            fx[f"G_{k}"] = G
        if Uk is not None and k in keep_u:
            fx[f"U_{k}"] = Uk
    fx["G_sites"] = sampled_sites(V)
    if trace_kmax:
        d = trace(harness, path, ns, nt, trace_kmax)
        fx["trace_k"] = d[:, 0].astype(int)
        fx["trace_obs"] = d[:, 1:]
    for est, tol in full_runs:
        rc, obs, G = parse_full(run(harness, "full", path, ns, nt, 10, est, tol), V)
        fx[f"full_est{est}_rc"] = rc
        fx[f"full_est{est}_obs"] = obs
        if est == 1000:
            fx[f"full_est{est}_G"] = G
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **fx)
    print(name, "done", {k: fx[k] for k in fx if k.startswith("full") and k.endswith("rc")})


D_ZZ1 = np.diag([-1, -1, 1]).astype(np.complex128)
D_1ZZ = np.diag([1, -1, -1]).astype(np.complex128)
P_II = np.diag([1j, 1j, -1])


def degenerate_config(ns: int = 4, nt: int = 8, seed: int = 31, eps: float = 0.3):
    """A smooth configuration (gfc.generate_smooth) whose spatial links on
    timeslices 0-2 are replaced by exact diagonal SU(3) patterns (temporal
    links and timeslices 3.. stay random).  w(n) only couples links of one
    timeslice, so each pattern keeps its slice in one branch of the local
    update (src/gauge_fixing.c:537-596, 656-708) at every half-sweep:

      t=0  U_i(n) = 1 for even n_i and diag(-1,-1,1) for odd n_i (i = Z, Y, X):
           w = diag(0,0,6), det w = 0, so inverse3x3 gives the null matrix and
           g_LA = 0 (:586-593); the OR update is Proj((1 - omega) 1), which is
           diag(-1,-1,1) for omega 1.5 (it flips every spatial link of the
           slice, and the pattern repeats on the other parity) and a failed
           projection -> identity for omega 1.0 (:701-705);
      t=1  U_Z = 1, U_Y = diag(-1,-1,1), U_X = diag(1,-1,-1) everywhere:
           w = diag(2,-2,2); the R and T Cabibbo-Marinari vectors are null and
           those subgroup steps are skipped (src/SU3_ops.c:829-839), S is the
           identity, g = 1;
      t=2  U_Z(n) = diag(i,i,-1) for even z and 1 for odd z, U_Y and U_X as on
           t=0: w = diag(1 +- i, 1 +- i, 4), det w = +-8i, so inverse3x3's
           double (real) return is 0 and no hit runs (:569); g = w w^-1 = 1.

    With omega 1.5 or 1.0 every value on these slices stays exact (entries
    0, +-1, +-i, 1/2), so reference and engine must agree on them bit for
    bit."""
    geom = gfc.Geometry(ns, nt)
    U = gfc.generate_smooth(geom, seed, eps).reshape(nt, ns, ns, ns, 4, 3, 3)
    one = np.eye(3, dtype=np.complex128)
    idx = np.arange(ns)
    for t in range(3):
        for mu, axis in ((1, 0), (2, 1), (3, 2)):  # mu Z, Y, X -> array axes z, y, x
            odd = (idx % 2 == 1)
            shape = [1, 1, 1]
            shape[axis] = ns
            odd = odd.reshape(shape)
            if t == 0 or (t == 2 and mu != 1):
                pat = np.where(odd[..., None, None], D_ZZ1, one)
            elif t == 1:
                pat = np.broadcast_to({1: one, 2: D_ZZ1, 3: D_1ZZ}[mu], (1, 1, 1, 3, 3))
            else:  # t == 2, mu == Z
                pat = np.where(odd[..., None, None], one, P_II)
            U[t, :, :, :, mu] = np.broadcast_to(pat, (ns, ns, ns, 3, 3))
    return geom, np.ascontiguousarray(U.reshape(geom.volume, 4, 3, 3))


def degenerate(tmp: str):
    geom, U = degenerate_config()
    ns, nt, V = geom.ns, geom.nt, geom.volume
    path = cfg_file(tmp, geom, U, 8, False)
    fx = {"ns": ns, "nt": nt, "seed": 31, "eps": 0.3, "config_sha256": sha(U), "R": 10,
          "hits": 2}
    for omega in ("1.5", "1.0"):
        os.environ["REF_OMEGA"] = omega
        tag = omega.replace(".", "")
        t = parse_traj(run("ref_golden_le64", "traj", path, ns, nt, 10, *KEEP_G), V, KEEP_G)
        fx[f"om{tag}_traj_k"] = np.array([r[0] for r in t])
        fx[f"om{tag}_traj_rc"] = np.array([r[1] for r in t])
        for k, _, _, G, Uk in t:
            fx[f"om{tag}_G_{k}"] = G
            fx[f"om{tag}_U_{k}"] = Uk
        d = trace("ref_golden_le64", path, ns, nt, 200)
        fx[f"om{tag}_trace_obs"] = d[:, 1:]
    del os.environ["REF_OMEGA"]
    np.savez_compressed(os.path.join(OUT, "lat_degen_4x8.npz"), **fx)
    print("lat_degen_4x8 done")


def prims():
    n = 256
    d = run("ref_golden_le64", "prims", n, 12345)
    per = 18 + 2 + 18 + 1 + 18 + 1 + 18 + 1 + 18 + 54 + 9 + 18
    assert d.size == n * per
    d = d.reshape(n, per)
    c = lambda a: np.ascontiguousarray(a).view(np.complex128).reshape(n, 3, 3)
    i = 0
    fx = {}
    fx["A"] = c(d[:, i:i + 18]); i += 18
    fx["det"] = d[:, i] + 1j * d[:, i + 1]; i += 2
    fx["inverse"] = c(d[:, i:i + 18]); i += 18
    fx["proj_status"] = d[:, i]; i += 1
    fx["proj"] = c(d[:, i:i + 18]); i += 18
    fx["pow195_status"] = d[:, i]; i += 1
    fx["pow195"] = c(d[:, i:i + 18]); i += 18
    fx["pow1_status"] = d[:, i]; i += 1
    fx["pow1"] = c(d[:, i:i + 18]); i += 18
    for s in "RST":
        fx[f"cm_{s}"] = c(d[:, i:i + 18]); i += 18
    fx["decomp"] = d[:, i:i + 9]; i += 9
    fx["prod_next"] = c(d[:, i:i + 18]); i += 18
    assert i == per
    np.savez_compressed(os.path.join(OUT, "prims.npz"), **fx)
    print("prims done")


def io(tmp: str):
    geom = gfc.Geometry(4, 8)
    U = gfc.generate_smooth(geom, 7, 0.3)
    fx = {"config_sha256": sha(U)}
    for tag, prec, swap, harness in [("be64", 8, True, "ref_golden_be64"),
                                     ("le64", 8, False, "ref_golden_le64"),
                                     ("le32", 4, False, "ref_golden_le32")]:
        path = cfg_file(tmp, geom, U, prec, swap)
        d = run(harness, "load", path, 4, 8)
        fx[f"{tag}_rc"] = int(d[0])
        fx[f"{tag}_sha256"] = sha(d[1:])
        fx[f"{tag}_file_sha256"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        raw = open(path, "rb").read()
        short = os.path.join(tmp, f"short_{tag}.bin")
        open(short, "wb").write(raw[:-1])
        fx[f"{tag}_short_rc"] = int(run(harness, "load", short, 4, 8)[0])
        long = os.path.join(tmp, f"long_{tag}.bin")
        open(long, "wb").write(raw + b"\0")
        fx[f"{tag}_long_rc"] = int(run(harness, "load", long, 4, 8)[0])
        fx[f"{tag}_missing_rc"] = int(run(harness, "load", os.path.join(tmp, "nope"), 4, 8)[0])
    np.savez_compressed(os.path.join(OUT, "io.npz"), **fx)
    print("io done")


PAR_FILES = {
    "sample": "#generic parameters\ngfix_proxy:\t\t\t\t\t            e2    \n"
              "tolerance:\t\t\t\t\t            1.0E-16\nmax_sweeps_to_fix:        \t\t\t50000\n"
              "estimate_sweeps_to_gf_progress:     1000\n\n#   parameters specific to "
              "overrelaxation \nomega_OR:                       1.97\nhits:          "
              "                 2\t\n\n#reunitarization parameter\n"
              "sweeps_to_reunitarization:  \t\t250\n",
    "theta": "gfix_proxy: theta\ntolerance: 1e-12\nomega_OR: 1.5\nhits: 3.7\n"
             "sweeps_to_reunitarization: 10\nestimate_sweeps_to_gf_progress: 0\n",
    "readme_keys": "max_hits: 5\nestimate_sweeps_to_gfixprogress: 10\ntolerance: 1e-14\n",
    "bad_omega": "omega_OR: 2.0\n",
    "bad_tol": "tolerance: -1\n",
    "bad_proxy": "gfix_proxy: e3\ntolerance: 1e-10\n",
    "zero_reunit": "sweeps_to_reunitarization: 0\n",
    "junk": "no colon here\n::\nkey_only:\n: value_only\nomega_OR:1.2:3\n",
    "spaces": "  o m e g a _ O R  :  1 . 9 \n t o l e r a n c e : 1e-15\n",
}


def params(tmp: str):
    res = {}
    for name, text in PAR_FILES.items():
        path = os.path.join(tmp, f"{name}.par")
        open(path, "w").write(text)
        d = run("ref_golden_le64", "params", path)
        res[name] = {"text": text, "tolerance": d[0], "proxy": int(d[1]),
                     "max_sweeps_to_fix": int(d[2]), "estimate": int(d[3]),
                     "reunit": int(d[4]), "error": int(d[5]), "omega": d[6], "hits": int(d[7])}
    d = run("ref_golden_le64", "params", os.path.join(tmp, "does_not_exist.par"))
    res["missing_file"] = {"text": None, "tolerance": d[0], "proxy": int(d[1]),
                           "max_sweeps_to_fix": int(d[2]), "estimate": int(d[3]),
                           "reunit": int(d[4]), "error": int(d[5]), "omega": d[6],
                           "hits": int(d[7])}
    json.dump(res, open(os.path.join(OUT, "params.json"), "w"), indent=1)
    print("params done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-large", action="store_true")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True,
                   stdout=subprocess.DEVNULL)
    only = set(a.only.split(",")) - {""}
    with tempfile.TemporaryDirectory() as tmp:
        if not only or "prims" in only:
            prims()
        if not only or "params" in only:
            params(tmp)
        if not only or "io" in only:
            io(tmp)
        if not only or "4x8" in only:
            lattice(tmp, "lat_4x8", 4, 8, 0.25, 1, range(1, 201), KEEP_G, "ref_golden_le64")
        if not only or "8x16" in only:
            lattice(tmp, "lat_8x16", 8, 16, 0.25, 2, range(1, 201), [200], "ref_golden_le64",
                    keep_u=())
        # the smallest lattice (the +mu and -mu neighbours coincide) and a
        # spatial extent that is not a multiple of 4
        if not only or "2x2" in only:
            lattice(tmp, "lat_2x2", 2, 2, 0.3, 11, range(1, 201), KEEP_G, "ref_golden_le64")
        if not only or "6x4" in only:
            lattice(tmp, "lat_6x4", 6, 4, 0.3, 12, range(1, 201), KEEP_G, "ref_golden_le64")
        if not only or "degen" in only:
            degenerate(tmp)
        if (not only and not a.skip_large) or "32x64" in only:
            lattice(tmp, "lat_32x64", 32, 64, 0.3, 3, [200], [200], "ref_golden_le64_mt",
                    full_runs=((1000, 1e-16),), keep_u=(), trace_kmax=200)


if __name__ == "__main__":
    main()
