"""pn1_force_range: the accelerations of one i shard (a run of 256-star
blocks) against every particle, as one rank of a run whose stars are split
across GPUs would compute them.  The shard's pair sums run in the same
blocks, j chunks and order as in the full evaluation, so its rows equal
pn1_force's bit for bit, in both forms and for any j split."""
import numpy as np
import pytest

from pn_amd import lib as pn

pytestmark = pytest.mark.gpu

RANGES = [(1, 257), (257, 1025), (1793, 2000), (1025, 2000), (1, 2000)]


@pytest.fixture(scope="module")
def ctx():
    c = pn.Context(capacity=2000)
    yield c
    c.close()


@pytest.mark.parametrize("mode", [pn.FP64, pn.MIXED])
@pytest.mark.parametrize("jsplit", [0, 3, 4])
def test_force_range_equals_force_rows(ctx, mode, jsplit):
    x, v, m = pn.ic_plummer(1999, 41, a=1.0, m_bh=0.05)
    ctx.set_tuning(jsplit, 0, -1)
    ctx.set_params(1e-3, 50.0, 2.0**-10, mode)
    ctx.upload(x, v, m)
    full = ctx.force()
    for i0, i1 in RANGES:
        part = ctx.force_range(i0, i1)
        assert np.array_equal(part, full[i0:i1]), (mode, jsplit, i0, i1)
    ctx.set_tuning(0, 0, -1)


def test_force_range_rejects_bad_ranges(ctx):
    x, v, m = pn.ic_plummer(1999, 42, a=1.0, m_bh=0.05)
    ctx.set_params(1e-3, 50.0, 2.0**-10, pn.FP64)
    ctx.upload(x, v, m)
    for i0, i1 in [(0, 257), (2, 257), (1, 300), (257, 257), (1, 2001)]:
        with pytest.raises(pn.Pn1Error, match="block boundaries"):
            ctx.force_range(i0, i1)
