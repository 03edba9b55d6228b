"""Files streamed between disk and HBM (gfc_job_load_config,
gfc_job_write_config, gfc_job_write_gauge_transf).

The device path must read and write exactly the bytes of the host loader and
writers (host_io.c, which follow src/fields_io.c), with the same return codes:
FP64 and FP32, big- and little-endian, over files of several 16 MiB pinned
blocks with a ragged last block.
"""
import os

import numpy as np
import pytest

import gfc

pytestmark = pytest.mark.gpu

# 16^3 x 32: an FP64 configuration is 75.5 MB (4.5 pinned blocks), FP32 37.7 MB
GEOM = gfc.Geometry(16, 32)


@pytest.fixture(scope="module")
def U():
    return gfc.generate_smooth(GEOM, 31, 0.3)


@pytest.mark.parametrize("prec,swap", [(8, True), (8, False), (4, True), (4, False)])
def test_load_matches_host_loader(ctx, tmp_path, U, prec, swap):
    path = str(tmp_path / "cfg.bin")
    assert gfc.write_config(path, GEOM, U, prec, swap) == 0
    rc_h, Uh = gfc.load_config(path, GEOM, prec, swap)
    assert rc_h == 0
    with gfc.Job(ctx, GEOM) as job:
        assert job.load_config(path, prec, swap) == 0
        Ud = job.download_config()
    assert np.array_equal(Ud.view(np.uint64), Uh.view(np.uint64))
    if prec == 8:
        assert np.array_equal(Ud, U)


def test_load_error_codes(ctx, tmp_path, U):
    path = str(tmp_path / "cfg.bin")
    assert gfc.write_config(path, GEOM, U, 8, True) == 0
    with gfc.Job(ctx, GEOM) as job:
        assert job.load_config(str(tmp_path / "missing.bin")) == -1
        assert job.load_config(path, 5) == -4
        size = os.path.getsize(path)
        with open(path, "r+b") as f:
            f.truncate(size - 8)
        assert job.load_config(path) == -2
        assert gfc.load_config(path, GEOM)[0] == -2
        with open(path, "r+b") as f:
            f.truncate(size + 1)
        assert job.load_config(path) == -3
        assert gfc.load_config(path, GEOM)[0] == -3


@pytest.mark.parametrize("prec,swap", [(8, True), (8, False), (4, True)])
def test_write_config_matches_host_writer(ctx, tmp_path, U, prec, swap):
    dev, host = str(tmp_path / "dev.bin"), str(tmp_path / "host.bin")
    with gfc.Job(ctx, GEOM) as job:
        job.upload_config(U)
        assert job.write_config(dev, prec, swap) == 0
    assert gfc.write_config(host, GEOM, U, prec, swap) == 0
    assert open(dev, "rb").read() == open(host, "rb").read()


def test_write_gauge_transf_matches_host_writer(ctx, tmp_path, U):
    dev, host = str(tmp_path / "g_dev.bin"), str(tmp_path / "g_host.bin")
    p = gfc.Params(tolerance=1e-300, max_sweeps_to_fix=20, estimate_sweeps_to_gf_progress=10**6)
    with gfc.Job(ctx, GEOM) as job:
        job.upload_config(U)
        job.reunitarize()
        rc, rep = job.fix(p)
        assert rc == gfc.ENOTCONVERGED and rep.sweeps == 20
        assert job.write_gauge_transf(dev) == 0
        G = job.download_gauge()
    assert gfc.write_gauge_transf(host, GEOM, G) == 0
    data = open(dev, "rb").read()
    assert len(data) == GEOM.volume * 144
    assert data == open(host, "rb").read()
    rc, G2 = gfc.load_gauge_transf(dev, GEOM)
    assert rc == 0 and np.array_equal(G2, G)
