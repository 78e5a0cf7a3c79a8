"""pn1_force_noalloc, the force entry point that allocates nothing (its
intermediates live in the library's fixed per-GPU scratch), against
pn1_force with the same j split: the same bits, on two calls in a row, the
second with a smaller n (the scratch then still holds the first call's
contents beyond the second call's extent).

Device buffers come from hipMalloc of the HIP runtime libpn1 itself is
linked to (test code allocates the caller's inputs and outputs; the entry
point allocates nothing).
"""
import ctypes

import numpy as np
import pytest

from pn_amd import lib as pn

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    pn.load()
    L = ctypes.CDLL("libamdhip64.so.7")     # the runtime libpn1 already loaded
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class Dev:
    """A device buffer holding a float64 array's bytes."""

    def __init__(self, L, nbytes):
        self.L, self.p = L, ctypes.c_void_p()
        assert L.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        assert self.L.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.empty(shape)
        assert out.nbytes == self.nbytes
        assert self.L.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.L.hipFree(self.p)


def noalloc_force(L, x, v, m, eps, c):
    n = len(m)
    st, md, acc = Dev(L, 48 * n), Dev(L, 8 * n), Dev(L, 24 * n)
    try:
        st.put(np.concatenate([x.T, v.T]))          # 6 rows of n: x, y, z, vx, vy, vz
        md.put(m)
        acc.put(np.full((3, n), np.nan))
        pn.force_noalloc(st.p.value, md.p.value, n, eps, c, acc.p.value, 0)
        assert L.hipDeviceSynchronize() == 0
        return acc.get((3, n)).T.copy()
    finally:
        for b in (st, md, acc):
            b.free()


def pn1_force(x, v, m, eps, c):
    with pn.Context(len(m)) as ctx:
        ctx.set_params(eps, c, 2.0**-10, pn.FP64)
        ctx.set_tuning(pn.noalloc_split(len(m)), 0, -1)
        ctx.upload(x, v, m)
        return ctx.force()


def test_two_calls_equal_pn1_force():
    L = hip()
    eps, c = 1e-3, 50.0
    cases = [pn.ic_plummer(pn.NOALLOC_MAX_N - 1, 41, a=1.0, m_bh=0.05),
             pn.ic_uniform_sphere(1000, 42, R=1.0, Q=0.5, m_bh=0.01, eps=eps)]
    for x, v, m in cases:
        print(f"case: pn1_force_noalloc n={len(m)} (split {pn.noalloc_split(len(m))})",
              flush=True)
        a = noalloc_force(L, x, v, m, eps, c)
        ref = pn1_force(x, v, m, eps, c)
        assert np.isfinite(a).all()
        assert np.array_equal(a, ref), np.abs(a - ref).max()


def test_small_n():
    """n = 1 (the BH alone) and n = 2 (one star: one block, split 1)."""
    L = hip()
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, -0.2]])
    v = np.array([[0.0, 0.0, 0.0], [0.1, 0.2, 0.05]])
    m = np.array([0.05, 1.0])
    for n in (1, 2):
        print(f"case: pn1_force_noalloc n={n}", flush=True)
        a = noalloc_force(L, x[:n], v[:n], m[:n], 1e-3, 50.0)
        if n == 1:
            assert np.array_equal(a, np.zeros((1, 3)))
        else:
            assert np.array_equal(a, pn1_force(x, v, m, 1e-3, 50.0))


def test_rejects_bad_arguments_before_launching():
    L = hip()
    n = pn.NOALLOC_MAX_N + 1
    st, md, acc = Dev(L, 48 * n), Dev(L, 8 * n), Dev(L, 24 * n)
    try:
        for nn, eps, c in ((n, 1e-3, 50.0), (0, 1e-3, 50.0), (100, 0.0, 50.0),
                           (100, 1e-3, -1.0)):
            with pytest.raises(pn.Pn1Error) as ei:
                pn.force_noalloc(st.p.value, md.p.value, nn, eps, c, acc.p.value, 0)
            assert ei.value.code == pn.PN1_EINVAL
        assert L.hipDeviceSynchronize() == 0
    finally:
        for b in (st, md, acc):
            b.free()
