"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and, where it was
built, of the compiled reference (oracle/_ref/libls9ref.so).  Test
infrastructure: the checker, never the thing measured."""
import contextlib
import ctypes
import os
import subprocess
import tempfile
import threading

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libls9ref.so")

_D = ctypes.POINTER(ctypes.c_double)
_dbl, _int = ctypes.c_double, ctypes.c_int


def ptr(a):
    return a.ctypes.data_as(_D)


_orc = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, os.path.basename(ORACLE_LIB)],
                           check=True)
        L = ctypes.CDLL(ORACLE_LIB)
        L.orc_force.argtypes = [_D, _D, _D, _dbl, _dbl, _D, _int]
        L.orc_force_omp.argtypes = [_D, _D, _D, _dbl, _dbl, _D, _int, _int]
        L.orc_force_omp.restype = _int
        L.orc_force_rows.argtypes = [_D, _D, _D, _dbl, _dbl, ctypes.POINTER(_int), _int,
                                     _D, _int]
        L.orc_energy_terms.argtypes = [_D, _D, _D, _int, _dbl, _dbl, _D]
        L.orc_initial_energy.argtypes = [_D, _D, _D, _int, _D, _dbl, _dbl]
        L.orc_energy.argtypes = [_dbl, _D, _D, _D, _int, _dbl, _dbl, _dbl]
        L.orc_q1_steps.argtypes = [_D, _D, _D, _int, _dbl, _dbl, _dbl, _int]
        L.orc_force_cs.argtypes = [_D, _D, _D, _dbl, _dbl, _D, _int]
        L.orc_q1_steps_cs.argtypes = [_D, _D, _D, _int, _dbl, _dbl, _dbl, _int]
        _orc = L
    return _orc


def have_ref():
    return os.path.exists(REF_LIB)


_ref = None


def ref():
    """The compiled reference (force/energy/initial_energy with c = 1)."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_LIB)
        L.force.argtypes = [_D, _D, _D, _dbl, _D, _D, _int]
        L.energy.argtypes = [_dbl, _D, _D, _D, _int, _dbl, _dbl]
        L.initial_energy.argtypes = [_D, _D, _D, _int, _D, _dbl]
        _ref = L
    return _ref


def big_stack(fn):
    """Runs fn in a thread with a 256 MiB stack: the reference keeps
    NMAX-sized arrays on the stack (6.5 MB in force())."""
    out = {}
    old = threading.stack_size(256 << 20)
    try:
        def run():
            try:
                out["r"] = fn()
            except BaseException as e:  # pragma: no cover
                out["e"] = e
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("r")


_libc = ctypes.CDLL(None)


@contextlib.contextmanager
def captured_stdout():
    """Captures what C code writes to file descriptor 1."""
    buf = {}
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        _libc.fflush(None)
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            yield buf
        finally:
            _libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
            tmp.seek(0)
            buf["text"] = tmp.read().decode()


def _arr(x, v, m):
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float64)
    return x, v, m


# ---------------------------------------------------------------- oracle calls
def orc_force(x, v, m, eps, c=1.0):
    x, v, m = _arr(x, v, m)
    a = np.zeros_like(x)
    oracle().orc_force(ptr(x), ptr(v), ptr(m), eps, c, ptr(a), m.shape[0])
    return a


def orc_force_omp(x, v, m, eps, c, threads):
    """force() over OpenMP threads: (a, number of threads that ran)."""
    x, v, m = _arr(x, v, m)
    a = np.zeros_like(x)
    ran = oracle().orc_force_omp(ptr(x), ptr(v), ptr(m), eps, c, ptr(a), m.shape[0], threads)
    return a, ran


def orc_force_rows(x, v, m, eps, c, rows):
    """Rows `rows` (stars) of orc_force, each O(n): bitwise those rows of the
    full call."""
    x, v, m = _arr(x, v, m)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((rows.shape[0], 3))
    oracle().orc_force_rows(ptr(x), ptr(v), ptr(m), eps, c,
                            rows.ctypes.data_as(ctypes.POINTER(_int)), rows.shape[0],
                            ptr(out), m.shape[0])
    return out


def orc_energy_terms(x, v, m, eps, c=1.0):
    x, v, m = _arr(x, v, m)
    e = np.zeros(5)
    oracle().orc_energy_terms(ptr(x), ptr(v), ptr(m), m.shape[0], eps, c, ptr(e))
    return e


def orc_initial_energy_text(x, v, m, eps, c=1.0):
    x, v, m = _arr(x, v, m)
    e = _dbl()
    with captured_stdout() as out:
        oracle().orc_initial_energy(ptr(x), ptr(v), ptr(m), m.shape[0], ctypes.byref(e), eps, c)
    return e.value, out["text"]


def orc_energy_text(t, x, v, m, init_ene, eps, c=1.0):
    x, v, m = _arr(x, v, m)
    with captured_stdout() as out:
        oracle().orc_energy(t, ptr(x), ptr(v), ptr(m), m.shape[0], init_ene, eps, c)
    return out["text"]


def orc_q1_steps(x, v, m, eps, c, dt, steps):
    x, v, m = _arr(x, v, m)
    x, v = x.copy(), v.copy()
    oracle().orc_q1_steps(ptr(x), ptr(v), ptr(m), m.shape[0], eps, c, dt, steps)
    return x, v


# ---------------------------------------------------------------- complex step
CS_H = 1e-30     # imaginary step: far below any rounding of the real parts


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _cptr(a):
    return a.ctypes.data_as(_D)


def orc_force_cs(x, v, m, eps, c=1.0):
    """The oracle's force on complex inputs (complex arrays in and out)."""
    x, v, m = _cplx(x), _cplx(v), _cplx(m)
    a = np.zeros_like(x)
    oracle().orc_force_cs(_cptr(x), _cptr(v), _cptr(m), eps, c, _cptr(a), m.shape[0])
    return a


def orc_q1_steps_cs(x, v, m, eps, c, dt, steps):
    x, v, m = _cplx(x).copy(), _cplx(v).copy(), _cplx(m)
    oracle().orc_q1_steps_cs(_cptr(x), _cptr(v), _cptr(m), m.shape[0], eps, c, dt, steps)
    return x, v


def _perturbed(x, v, m, coord):
    """Complex copies of (x, v, m) with i*CS_H added to one input:
    coord = ("x" | "v", particle, axis) or ("m0",)."""
    xc, vc, mc = _cplx(x).copy(), _cplx(v).copy(), _cplx(m).copy()
    if coord[0] == "x":
        xc[coord[1], coord[2]] += 1j * CS_H
    elif coord[0] == "v":
        vc[coord[1], coord[2]] += 1j * CS_H
    else:
        mc[0] += 1j * CS_H
    return xc, vc, mc


def cs_force_vjp(x, v, m, u, eps, c, coords):
    """d(sum_p u_p . a_p)/d(coord) for each coord, by complex step on the
    oracle's force."""
    out = []
    for co in coords:
        a = orc_force_cs(*_perturbed(x, v, m, co), eps, c)
        out.append(float(np.sum(u * a.imag)) / CS_H)
    return np.array(out)


def cs_trajectory_grad(x, v, m, eps, c, dt, steps, loss_grad, coords):
    """d loss(final state)/d(coord) for each coord: complex step through the
    oracle's stage driver, contracted with loss_grad = (dL/dx_T, dL/dv_T)
    evaluated at the unperturbed final state (exact for the first-order
    term, which is all the imaginary part carries)."""
    gx, gv = loss_grad
    out = []
    for co in coords:
        xt, vt = orc_q1_steps_cs(*_perturbed(x, v, m, co), eps, c, dt, steps)
        out.append(float(np.sum(gx * xt.imag) + np.sum(gv * vt.imag)) / CS_H)
    return np.array(out)


# ---------------------------------------------------------------- reference calls
def ref_force(x, v, m, eps):
    x, v, m = _arr(x, v, m)
    a = np.zeros_like(x)
    pot = np.zeros(m.shape[0])
    big_stack(lambda: ref().force(ptr(x), ptr(v), ptr(m), eps, ptr(a), ptr(pot), m.shape[0]))
    return a


def ref_initial_energy_text(x, v, m, eps):
    x, v, m = _arr(x, v, m)
    e = _dbl()
    with captured_stdout() as out:
        big_stack(lambda: ref().initial_energy(ptr(x), ptr(v), ptr(m), m.shape[0],
                                               ctypes.byref(e), eps))
    return e.value, out["text"]


def ref_energy_text(t, x, v, m, init_ene, eps):
    x, v, m = _arr(x, v, m)
    with captured_stdout() as out:
        big_stack(lambda: ref().energy(t, ptr(x), ptr(v), ptr(m), m.shape[0], init_ene, eps))
    return out["text"]
