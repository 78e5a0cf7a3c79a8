"""Refused device allocations of the pair partials leave the library usable
and never lead to a launch on a buffer that does not cover it.

  - pn1_set_tuning with a j split whose partials cannot be allocated returns
    PN1_ENOMEM and keeps the shape in use before; the next FP64 evaluation
    gives the same bits as before the call (mixed: the same after a new
    upload; without one, within 1e-6, its thresholds come from the last
    evaluation).
  - pn1_ens_upload whose work-item tables need more partials than the device
    holds returns PN1_ENOMEM and withdraws the ensemble (the next force is
    refused as "no ensemble uploaded", nothing is launched); a later upload
    that fits works.
The sizes are chosen so that the refused requests exceed the 288 GB of an
MI355X by far (8.8 TB and 550 GB).
"""
import numpy as np
import pytest

from conftest import rel_err
from pn_amd import lib as pn

pytestmark = pytest.mark.gpu


def test_refused_split_keeps_the_shape():
    x, v, m = pn.ic_plummer(4096, 31, a=1.0, m_bh=0.02)
    # capacity 2^24: a split of 4096 chunks needs 4096 x 16 x 2^24 doubles
    c = pn.Context(capacity=1 << 24)
    try:
        for mode in (pn.FP64, pn.MIXED):
            print(f"case: refused j split, mode {mode}", flush=True)
            c.set_params(1e-3, 100.0, 2.0**-10, mode)
            c.set_tuning(0, 0, -1)
            c.upload(x, v, m)
            shape = c.get_tuning()
            a0 = c.force()
            with pytest.raises(pn.Pn1Error) as ei:
                c.set_tuning(4096, 0, -1)
            assert ei.value.code == pn.PN1_ENOMEM
            assert c.get_tuning() == shape
            a1 = c.force()
            if mode == pn.FP64:
                assert np.array_equal(a1, a0)
            else:                   # thresholds now from a0: the same to 1e-6
                assert rel_err(a1, a0).max() <= 1e-6
                c.upload(x, v, m)
                assert np.array_equal(c.force(), a0)
            c.step(1)
            xs, _ = c.download()
            assert np.isfinite(xs).all()
    finally:
        c.close()


def test_refused_ensemble_tables_withdraw_the_state():
    big = pn.ic_plummer((1 << 20) - 1, 32, a=1.0, m_bh=0.02)     # 2^20 particles
    small = [pn.ic_plummer(300, 33 + k, a=1.0, m_bh=0.02) for k in range(3)]
    e = pn.Ensemble(capacity=1 << 20, max_clusters=4)
    try:
        e.set_params(1e-3, 100.0, 2.0**-10, pn.FP64)
        e.upload(small)
        a0 = e.force()
        # chunks of 256: 4096 x 4096 items, 2^32 partial slots of 128 bytes
        print("case: ensemble tables that cannot be allocated", flush=True)
        e.set_tuning(256, 0, -1)
        with pytest.raises(pn.Pn1Error) as ei:
            e.upload([big])
        assert ei.value.code == pn.PN1_ENOMEM
        with pytest.raises(pn.Pn1Error, match="no ensemble uploaded"):
            e.force()
        e.set_tuning(0, 0, -1)
        e.upload(small)
        a1 = e.force()
        for p, q in zip(a0, a1):
            assert np.array_equal(p, q)
    finally:
        e.close()
