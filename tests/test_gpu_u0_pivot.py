"""FP64 u0 storage: five numbers per link (row 0, and row 1 without the
element under row 0's pivot, rebuilt from the orthogonality of the rows).

Smooth configurations near the identity always pivot on column 0, so this
test feeds Haar-random links, whose row-0 pivot falls on every column, plus
links with exact ties between row-0 moduli (columns 0/1, 1/2, all three), and
checks the sweeps, which read u0 only through the five stored numbers, against
the CPU oracle (which keeps whole matrices) on the same input: e2 after each
of 20 half-sweeps to 1e-12 relative and G to 1e-10, the drop-in tolerances."""
import numpy as np
import pytest

import gfc
import oracle as orc

pytestmark = pytest.mark.gpu


def _haar_su3(rng, n):
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / 3.0))[:, None, None]


def _ties():
    c = 1 / np.sqrt(2)
    t01 = np.array([[c, c, 0], [-c, c, 0], [0, 0, 1]], complex)
    t12 = np.array([[0, c, c], [0, -c, c], [1, 0, 0]], complex)
    w = np.exp(2j * np.pi / 3)
    f = np.array([[1, 1, 1], [1, w, w * w], [1, w * w, w]]) / np.sqrt(3)
    f = f / np.linalg.det(f) ** (1.0 / 3.0)
    return [t01, t12, f]


def test_u0_pivots_and_ties_match_oracle(ctx):
    geom = gfc.Geometry(4, 8)
    V = geom.volume
    rng = np.random.default_rng(7)
    U = _haar_su3(rng, 4 * V).reshape(V, 4, 3, 3)
    ties = _ties()
    for k in range(0, V, 7):
        U[k, 1 + k % 3] = ties[k % 3]
    piv = np.argmax(np.abs(U[:, 1:, 0, :]) ** 2, axis=-1)
    assert all(np.count_nonzero(piv == c) > 100 for c in range(3))
    job = gfc.Job(ctx, geom)
    job.upload_config(U)
    assert job.reunitarize() == 0
    k = 20
    p = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=100000000,
                   max_sweeps_to_fix=k, sweeps_to_reunitarization=10)
    rc, rep, tr = job.fix(p, trace_len=k)
    assert rc == gfc.ENOTCONVERGED and rep.sweeps == k
    G = job.download_gauge()
    job.close()
    U1 = U.copy()
    og = orc.geom(4, 8)
    orc.reunitarize(U1)
    G1 = np.zeros((V, 3, 3), np.complex128)
    rc_o, tr_o = orc.gaugefix(U1, G1, og, orc.params(tolerance=1e-300, max_sweeps=k,
                                                     estimate=100000000, reunit_every=10),
                              trace_len=k)
    rel = np.max(np.abs(tr - tr_o) / tr_o)
    dG = np.max(np.abs(G - G1))
    print(f"Haar + ties 4^3x8: e2 over {k} sweeps max rel diff {rel:.2e}, max|dG| {dG:.2e}")
    assert rel < 1e-12
    assert dG < 1e-10
