"""The drop-in command line on the GPU: same outputs and exit codes as the
reference's gauge_fix_coulomb (gauge_fix_coulomb.c:67-211)."""
import os
import subprocess

import numpy as np
import pytest

import gfc
from conftest import golden

pytestmark = pytest.mark.gpu


def _run(args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([gfc.CLI_PATH, *args], cwd=cwd, env=e, capture_output=True,
                          text=True, timeout=120)


def _par(tmp_path, **kw):
    text = "".join(f"{k}: {v}\n" for k, v in kw.items())
    p = tmp_path / "parameters.par"
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("fmt", ["be64", "le64", "le32"])
def test_cli_matches_reference_run(tmp_path, fmt):
    f = golden("lat_4x8.npz")
    geom = gfc.Geometry(4, 8)
    U = gfc.generate_smooth(geom, int(f["seed"]), float(f["eps"]))
    prec, swap = {"be64": (8, True), "le64": (8, False), "le32": (4, False)}[fmt]
    cfg = str(tmp_path / f"cfg_{fmt}.bin")
    assert gfc.write_config(cfg, geom, U, prec, swap) == 0
    par = _par(tmp_path, sweeps_to_reunitarization=10)
    env = {"GFC_BYTE_SWAP_IN": str(int(swap)), "GFC_INPUT_PRECISION": str(prec)}
    r = _run([cfg, str(tmp_path / "gt.bin"), "4", "8", par], tmp_path, env)
    assert r.returncode == 0, r.stderr
    assert "Average trace of spatial plaquettes before gauge transformation" in r.stdout
    assert "Sweeps needed to gauge-fix config from file" in r.stdout
    rec = (tmp_path / "sweeps_to_gaugefix_4x8.txt").read_text()
    if fmt == "le32":
        # FP32 input is a different configuration (rounded); it must still converge
        assert rec.startswith(cfg + "\t")
        return
    assert rec == f"{cfg}\t{int(f['full_est1000_rc'])}\n"
    G = np.fromfile(tmp_path / "gt.bin", dtype="<c16").reshape(-1, 3, 3)
    assert G.shape[0] == geom.volume
    assert np.max(np.abs(G - f["full_est1000_G"])) < 1e-10


def test_cli_failures_write_no_gauge_file(tmp_path):
    geom = gfc.Geometry(4, 8)
    U = gfc.generate_smooth(geom, 1, 0.25)
    cfg = str(tmp_path / "cfg.bin")
    gfc.write_config(cfg, geom, U, 8, True)
    # wrong lattice size -> loadConfig fails (-3 trailing data) -> exit 1
    r = _run([cfg, str(tmp_path / "gt.bin"), "4", "4"], tmp_path)
    assert r.returncode == 1 and "Loading of file" in r.stderr
    assert not (tmp_path / "gt.bin").exists()
    # not converged within max sweeps -> exit 1, no G, no sweeps record
    par = _par(tmp_path, max_sweeps_to_fix=5)
    r = _run([cfg, str(tmp_path / "gt.bin"), "4", "8", par], tmp_path)
    assert r.returncode == 1 and "could not be gauge-fixed" in r.stderr
    assert not (tmp_path / "gt.bin").exists()
    assert not (tmp_path / "sweeps_to_gaugefix_4x8.txt").exists()
    # invalid parameters fall back to the defaults (a warning, not an exit)
    par = _par(tmp_path, omega_OR=2.5, sweeps_to_reunitarization=10)
    r = _run([cfg, str(tmp_path / "gt.bin"), "4", "8", par], tmp_path)
    assert r.returncode == 0 and "Initializing gauge-fixing parameters to default" in r.stderr
    assert (tmp_path / "gt.bin").stat().st_size == geom.volume * 144


def test_cli_mixed_precision(tmp_path):
    geom = gfc.Geometry(8, 16)
    U = gfc.generate_smooth(geom, 2, 0.25)
    cfg = str(tmp_path / "cfg.bin")
    gfc.write_config(cfg, geom, U, 8, True)
    par = _par(tmp_path, sweeps_to_reunitarization=10)
    r = _run([cfg, str(tmp_path / "gt.bin"), "8", "16", par], tmp_path,
             {"GFC_PRECISION": "mixed", "GFC_METRICS": "1"})
    assert r.returncode == 0, r.stderr
    import json
    m = json.loads(r.stdout.strip().splitlines()[-1])
    assert m["sweeps_fp32"] > 0 and m["residue_final"] <= 1e-16
