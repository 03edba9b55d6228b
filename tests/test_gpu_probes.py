"""The bandwidth probes behind bench.py's measured ceilings: a copy and a read
of a device buffer, timed with HIP events; bad arguments are refused."""
import math

import pytest

import gfc

pytestmark = pytest.mark.gpu


def test_copy_and_read_probes(ctx):
    for nbytes in (1 << 20, 64 << 20):
        for probe in (ctx.copy_probe, ctx.read_probe):
            s = probe(nbytes, 3)
            assert math.isfinite(s) and s > 0
    # a 16-byte buffer is the smallest accepted; below it the call fails
    assert ctx.read_probe(16, 1) > 0
    with pytest.raises(gfc.GfcError):
        ctx.read_probe(8, 1)
    with pytest.raises(gfc.GfcError):
        ctx.copy_probe(1 << 20, 0)
    with pytest.raises(gfc.GfcError):
        ctx.read_probe(1 << 20, (1 << 20) + 1)
