"""pn1_force_range: the accelerations of one i shard (a run of 256-star
blocks) against every particle, as one rank of a run whose stars are split
across GPUs would compute them.  With an explicit j split the shard's pair
sums run in the same blocks, j chunks and order as in the full evaluation,
so its rows equal pn1_force's bit for bit, in both forms.  With the
automatic plan a small shard gets more j chunks (a fuller grid), which only
reorders the sums."""
import numpy as np
import pytest

from conftest import rel_err
from pn_amd import lib as pn

pytestmark = pytest.mark.gpu

RANGES = [(1, 257), (257, 1025), (1793, 2000), (1025, 2000), (1, 2000)]


@pytest.fixture(scope="module")
def ctx():
    c = pn.Context(capacity=20001)
    yield c
    c.close()


@pytest.mark.parametrize("mode", [pn.FP64, pn.MIXED])
@pytest.mark.parametrize("jsplit", [1, 3, 4, 16])
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


@pytest.mark.parametrize("mode", [pn.FP64, pn.MIXED])
def test_force_range_own_split_agrees(ctx, mode):
    """n = 20000: the full plan splits j in 128 chunks, an 8-block shard in
    256; FP64 rows agree to 1e-13, mixed rows within the mixed form's 1e-6
    of FP64 each."""
    x, v, m = pn.ic_plummer(19999, 43, a=1.0, m_bh=0.05)
    ctx.set_tuning(0, 0, -1)
    ctx.set_params(1e-3, 50.0, 2.0**-10, mode)
    ctx.upload(x, v, m)
    full = ctx.force()
    for i0, i1 in [(1, 2049), (2049 + 256 * 40, 2049 + 256 * 48), (1 + 256 * 78, 20000)]:
        part = ctx.force_range(i0, i1)
        err = rel_err(part, full[i0:i1])
        bound = 1e-13 if mode == pn.FP64 else 2e-6
        assert err.max() <= bound, (mode, i0, i1, err.max())


def test_force_range_rejects_bad_ranges(ctx):
    x, v, m = pn.ic_plummer(1999, 42, a=1.0, m_bh=0.05)
    ctx.set_params(1e-3, 50.0, 2.0**-10, pn.FP64)
    ctx.upload(x, v, m)
    for i0, i1 in [(0, 257), (2, 257), (1, 300), (257, 257), (1, 2001)]:
        with pytest.raises(pn.Pn1Error, match="block boundaries"):
            ctx.force_range(i0, i1)
