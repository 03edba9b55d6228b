"""Object lifetimes in the Python binding: closing a context first closes the
jobs and servers made on it (the C objects use the context's streams when
they are destroyed), so a job that outlives its context, e.g. held by a
failed test's traceback, is never destroyed through a freed context."""
import pytest

import gfc

pytestmark = pytest.mark.gpu


def test_context_close_closes_its_jobs_and_servers():
    ctx = gfc.Context(0)
    job = gfc.Job(ctx, gfc.Geometry(4, 8))
    srv = gfc.Server(ctx, max_jobs=2)
    other = gfc.Job(ctx, gfc.Geometry(2, 2))
    other.close()
    ctx.close()
    assert not job._p and not srv._p and not ctx._p
    job.close()  # no-op once closed
    srv.close()
    del job, srv, other
    ctx.close()
