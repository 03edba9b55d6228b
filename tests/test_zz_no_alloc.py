"""gfc_config_observables_noalloc: the observables of a configuration already in
device memory, with the partial sums in the library's fixed device scratch
(no allocation).  Checked bit for bit against gfc_config_observables (host
array in, device job allocated per call) on two calls in a row, the second
on a smaller lattice, and the size limit of the scratch."""
import ctypes as C

import numpy as np
import pytest

import gfc

pytestmark = pytest.mark.gpu

_hip = C.CDLL("libamdhip64.so.7")  # the runtime libgfc.so is linked against
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_H2D = 1


def _to_device(a: np.ndarray) -> int:
    p = C.c_void_p()
    assert _hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert _hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, _H2D) == 0
    return p.value


def test_noalloc_matches_allocating_entry_point(ctx):
    for ns, nt, seed in ((32, 64, 21), (8, 16, 22)):
        geom = gfc.Geometry(ns, nt)
        U = gfc.generate_smooth(geom, seed, 0.3)
        ref = gfc.config_observables(ctx, geom, U)
        d = _to_device(U)
        try:
            got = gfc.config_observables_noalloc(ctx, geom, d)
        finally:
            _hip.hipFree(C.c_void_p(d))
        print(f"{ns}^3x{nt}: noalloc {got} allocating {ref}")
        assert got == ref


def test_noalloc_rejects_lattices_beyond_the_scratch(ctx):
    geom = gfc.Geometry(50, 90)  # 11 250 000 sites > 11 182 080
    p = C.c_void_p()
    nbytes = geom.volume * 4 * 18 * 8
    assert _hip.hipMalloc(C.byref(p), nbytes) == 0
    try:
        with pytest.raises(gfc.GfcError) as e:
            gfc.config_observables_noalloc(ctx, geom, p.value)
        assert e.value.code == gfc.EINVAL
    finally:
        _hip.hipFree(p)
