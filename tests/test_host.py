"""Host-side logic of libgfc (no GPU): parameters, schedule, geometry, SCIDAC
I/O, the sweeps record, the synthetic generator and the exported ABI.

Parameter-file and loader expectations are the reference's own results
(tests/golden/params.json and io.npz, written by the compiled reference).
"""
import hashlib
import json
import math
import os
import re

import numpy as np
import pytest

import gfc
import oracle as orc
from conftest import GOLDEN, ROOT, golden


def test_default_params_match_reference():
    d = gfc.Params.default()
    assert d.tolerance == 1e-16 and d.gfix_proxy == gfc.PROXY_E2
    assert d.max_sweeps_to_fix == 100000 and d.estimate_sweeps_to_gf_progress == 1000
    assert d.sweeps_to_reunitarization == 250 and d.hits == 2
    assert d.omega_OR == np.float32(1.95) and d.error == 0 and d.valid()
    ref = json.load(open(os.path.join(GOLDEN, "params.json")))["missing_file"]
    assert (d.tolerance, d.max_sweeps_to_fix, d.estimate_sweeps_to_gf_progress,
            d.sweeps_to_reunitarization, d.hits) == (
        ref["tolerance"], ref["max_sweeps_to_fix"], ref["estimate"], ref["reunit"], ref["hits"])


@pytest.mark.parametrize("name", ["sample", "theta", "readme_keys", "bad_omega", "bad_tol",
                                  "bad_proxy", "zero_reunit", "junk", "spaces"])
def test_parameter_file_parsing_matches_reference(tmp_path, name):
    ref = json.load(open(os.path.join(GOLDEN, "params.json")))[name]
    path = tmp_path / f"{name}.par"
    path.write_text(ref["text"])
    p = gfc.Params.from_file(str(path))
    assert p.tolerance == ref["tolerance"]
    assert p.gfix_proxy == ref["proxy"]
    assert p.max_sweeps_to_fix == ref["max_sweeps_to_fix"]
    assert p.estimate_sweeps_to_gf_progress == ref["estimate"]
    assert p.sweeps_to_reunitarization == ref["reunit"]
    assert p.error == ref["error"]
    assert np.float32(p.omega_OR) == np.float32(ref["omega"])
    assert p.hits == ref["hits"]


def test_missing_parameter_file_keeps_defaults():
    ref = json.load(open(os.path.join(GOLDEN, "params.json")))["missing_file"]
    p = gfc.Params.from_file("/nonexistent/parameters.par")
    assert p.error == ref["error"] == 0 and p.tolerance == ref["tolerance"]


@pytest.mark.parametrize("sweep,res,tol,est", [
    (0, 16.07, 1e-16, 1000), (0, 15.8, 1e-16, 1000), (120, 3.2e-5, 1e-16, 1000),
    (0, 16.0, 1e-16, 0), (50, 1e-9, 1e-14, 250), (0, 1e-17, 1e-16, 1000)])
def test_next_check_schedule(sweep, res, tol, est):
    p = gfc.Params(tolerance=tol, estimate_sweeps_to_gf_progress=est)
    op = orc.params(tolerance=tol, estimate=est)
    got = p.next_check(sweep, res)
    assert got == orc.next_check(sweep, res, op)
    if res < tol:
        assert got == sweep
    else:  # closed form of src/gauge_fixing.c:280-282
        assert got == sweep + 10 + int(est * (1.0 - math.log10(res) / math.log10(tol)))


def test_first_check_lands_at_1085_for_e2_16():
    p = gfc.Params()
    assert p.next_check(0, 16.07) == 1085
    assert p.next_check(0, 15.8) == 1084


def test_geometry_rules():
    chk = lambda ns, nt: gfc.lib.gfc_geom_check(__import__("ctypes").byref(gfc._Geom(ns, nt)))
    assert chk(4, 8) == 0 and chk(48, 96) == 0
    assert chk(0, 8) == gfc.EGEOMETRY and chk(4, -2) == gfc.EGEOMETRY
    assert chk(5, 8) == gfc.EGEOMETRY and chk(4, 7) == gfc.EGEOMETRY  # odd extents
    V = 48 ** 3 * 96
    # cfg 576 B + u0 3 x 96 B + G 96 B per site; mixed adds the FP32 copies
    assert gfc.footprint(gfc.Geometry(48, 96)) == 960 * V
    assert gfc.footprint(gfc.Geometry(48, 96), gfc.PREC_MIXED) == (960 + 192) * V
    assert gfc.footprint(gfc.Geometry(5, 8)) == -1


def test_loader_matches_reference(tmp_path):
    f = golden("io.npz")
    geom = gfc.Geometry(4, 8)
    U = gfc.generate_smooth(geom, 7, 0.3)
    assert hashlib.sha256(U.tobytes()).hexdigest() == str(f["config_sha256"])
    for tag, prec, swap in [("be64", 8, True), ("le64", 8, False), ("le32", 4, False)]:
        path = str(tmp_path / f"{tag}.bin")
        assert gfc.write_config(path, geom, U, prec, swap) == 0
        assert hashlib.sha256(open(path, "rb").read()).hexdigest() == str(
            f[f"{tag}_file_sha256"])
        rc, L = gfc.load_config(path, geom, prec, swap)
        assert rc == f[f"{tag}_rc"] == 0
        assert hashlib.sha256(L.view(np.float64).tobytes()).hexdigest() == str(
            f[f"{tag}_sha256"])
        raw = open(path, "rb").read()
        open(path + ".short", "wb").write(raw[:-1])
        assert gfc.load_config(path + ".short", geom, prec, swap)[0] == f[f"{tag}_short_rc"]
        open(path + ".long", "wb").write(raw + b"\0")
        assert gfc.load_config(path + ".long", geom, prec, swap)[0] == f[f"{tag}_long_rc"]
        assert gfc.load_config(str(tmp_path / "nope"), geom, prec, swap)[0] == f[
            f"{tag}_missing_rc"]


def test_byte_swap_is_an_involution(tmp_path):
    geom = gfc.Geometry(2, 2)
    U = gfc.generate_smooth(geom, 3, 0.2)
    p = str(tmp_path / "be.bin")
    gfc.write_config(p, geom, U, 8, True)
    raw = np.fromfile(p, dtype=">f8")
    assert np.array_equal(raw, U.view(np.float64).ravel())
    rc, L = gfc.load_config(p, geom, 8, True)
    assert rc == 0 and np.array_equal(L, U)


def test_gauge_transf_round_trip_and_size(tmp_path):
    geom = gfc.Geometry(4, 8)
    G = gfc.generate_smooth(geom, 9, 0.1)[:, 0].copy()
    p = str(tmp_path / "G.bin")
    assert gfc.write_gauge_transf(p, geom, G) == 0
    assert os.path.getsize(p) == geom.volume * 144  # V * 144 B (README.md:38 formula)
    assert np.array_equal(np.fromfile(p, "<c16").reshape(-1, 3, 3), G)
    rc, G2 = gfc.load_gauge_transf(p, geom)
    assert rc == 0 and np.array_equal(G2, G)


def test_sweeps_record_format(tmp_path):
    geom = gfc.Geometry(8, 16)
    assert gfc.write_sweeps_record(str(tmp_path), geom, "cfg_a.bin", 1085) == 0
    assert gfc.write_sweeps_record(str(tmp_path), geom, "cfg_b.bin", 770) == 0
    text = (tmp_path / "sweeps_to_gaugefix_8x16.txt").read_text()
    assert text == "cfg_a.bin\t1085\ncfg_b.bin\t770\n"


def test_generator_is_deterministic_and_su3():
    geom = gfc.Geometry(4, 4)
    a = gfc.generate_smooth(geom, 42, 0.3)
    b = gfc.generate_smooth(geom, 42, 0.3)
    c = gfc.generate_smooth(geom, 43, 0.3)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    uu = np.einsum("sxab,sxcb->sxac", a, a.conj())
    assert np.max(np.abs(uu - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(a) - 1)) < 1e-13


def test_generator_pure_gauge_has_unit_plaquette():
    """eps = 0 gives U = g(n) g(n+mu)^dagger: every plaquette is the identity."""
    geom = gfc.Geometry(4, 4)
    U = gfc.generate_smooth(geom, 5, 0.0)
    g = orc.geom(4, 4)
    assert abs(orc.plaquette(U, g, 0) - 1) < 1e-13 and abs(orc.plaquette(U, g, 1) - 1) < 1e-13


def test_abi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "gfc.h")).read()
    declared = set(re.findall(r"\b(gfc_[a-z0-9_]+)\s*\(", header))
    import subprocess
    syms = subprocess.run(["nm", "-D", "--defined-only", gfc.LIB_PATH], capture_output=True,
                          text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in syms.splitlines() if " T " in ln}
    missing = declared - exported
    assert not missing, missing
    assert set(gfc.EXPORTED) <= exported
    assert os.access(gfc.CLI_PATH, os.X_OK)
    assert b"gfc-mi355x" in gfc.lib.gfc_version()
