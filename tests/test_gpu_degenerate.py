"""The degenerate branches of the local update on the GPU, against the
compiled reference (tests/golden/lat_degen_4x8.npz; the configuration is
tests/golden/make_golden.py degenerate_config):

  timeslice 0: det w(n) = 0 -> g_LA = 0 (src/gauge_fixing.c:586-593) -> OR
               diag(-1,-1,1) (omega 1.5), or a null matrix whose projection
               fails -> identity (omega 1.0, :701-705);
  timeslice 1: null Cabibbo-Marinari vectors -> subgroup skipped
               (src/SU3_ops.c:829-839);
  timeslice 2: Re det w(n) = 0 with det w(n) != 0 -> no hits (:569).

On those slices every value is exact, so G must equal the reference's bit for
bit, in FP64 and in mixed precision (FP32 sweeps); elsewhere the usual
tolerances hold (G 1e-10, e2 after every sweep 1e-12 relative, U 1e-12)."""
import hashlib
import os
import sys

import numpy as np
import pytest

import gfc
from conftest import GOLDEN, golden

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu


def _setup(ctx, precision=gfc.PREC_FP64):
    f = golden("lat_degen_4x8.npz")
    geom, U = make_golden.degenerate_config()
    assert hashlib.sha256(U.tobytes()).hexdigest() == str(f["config_sha256"])
    job = gfc.Job(ctx, geom, precision)
    job.upload_config(U)
    assert job.reunitarize() == 0
    job.save_config()
    return f, geom, job


def _params(k, omega):
    return gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=100000000,
                      max_sweeps_to_fix=k, sweeps_to_reunitarization=10, omega_OR=omega)


@pytest.mark.parametrize("omega", ["1.5", "1.0"])
def test_degenerate_branches_fp64(ctx, omega):
    f, geom, job = _setup(ctx)
    tag = omega.replace(".", "")
    S3 = 3 * geom.ns ** 3
    for k in [1, 2, 3, 10, 11, 200]:
        job.restore_config()
        if k == 200:
            rc, rep, tr = job.fix(_params(k, float(omega)), trace_len=200)
        else:
            rc, rep = job.fix(_params(k, float(omega)))
        assert rc == gfc.ENOTCONVERGED and rep.sweeps == k
        G = job.download_gauge()
        Gref = f[f"om{tag}_G_{k}"]
        dG = np.max(np.abs(G - Gref))
        print(f"omega {omega} k={k}: max|dG| {dG:.2e}")
        assert np.array_equal(G[:S3], Gref[:S3])
        assert dG < 1e-10
    ref = f[f"om{tag}_trace_obs"][:, 0]
    rel = np.max(np.abs(tr - ref) / ref)
    print(f"omega {omega}: e2 over 200 sweeps, max rel diff {rel:.2e}")
    assert rel < 1e-12
    U = job.download_config()
    assert np.max(np.abs(U - f[f"om{tag}_U_200"])) < 1e-12
    job.close()


@pytest.mark.parametrize("omega", ["1.5", "1.0"])
def test_degenerate_branches_mixed(ctx, omega):
    """FP32 sweeps take the same branches: after 200 half-sweeps (all in the
    FP32 phase, then handed over) G on the degenerate slices is the
    reference's, bit for bit."""
    f, geom, job = _setup(ctx, gfc.PREC_MIXED)
    tag = omega.replace(".", "")
    S3 = 3 * geom.ns ** 3
    o = gfc.Options(precision=gfc.PREC_MIXED, mixed_switch=1e-300)
    rc, rep = job.fix(_params(200, float(omega)), o)
    print(f"omega {omega}: mixed, {rep.sweeps_fp32} of {rep.sweeps} sweeps in FP32")
    assert rc == gfc.ENOTCONVERGED and rep.sweeps == 200 and rep.sweeps_fp32 >= 50
    G = job.download_gauge()
    assert np.array_equal(G[:S3], f[f"om{tag}_G_200"][:S3])
    job.close()
