"""ctypes binding of libpn1 (include/pn1.h).

This is plumbing for tests, the bench and scripting: every computation runs
in libpn1.so (HIP kernels for gfx950).  There is no fallback: if the
library is not built, or no gfx950 device is present, calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpn1.so")

PN1_OK, PN1_EINVAL, PN1_ENOMEM, PN1_EDEVICE, PN1_ENODEV, PN1_EIO = 0, -1, -2, -3, -4, -5
FP64, MIXED = 0, 1
PN1_MAX_CAPACITY = 1 << 28
_STATUS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ENODEV", -5: "EIO"}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


class Pn1Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpn1 {_STATUS.get(code, code)}: {msg}")
        self.code = code


class Energy(ctypes.Structure):
    _fields_ = [("kin", ctypes.c_double), ("pot", ctypes.c_double),
                ("pot_pn", ctypes.c_double), ("pot_pn2", ctypes.c_double),
                ("total", ctypes.c_double)]

    def as_array(self):
        return np.array([self.kin, self.pot, self.pot_pn, self.pot_pn2, self.total])


_lib = None


def load():
    """Loads libpn1.so (once) and declares its prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m pn_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    sig = {
        "force": (None, [_D, _D, _D, ctypes.c_double, _D, _D, _I]),
        "energy": (None, [ctypes.c_double, _D, _D, _D, _I, ctypes.c_double, ctypes.c_double]),
        "initial_energy": (None, [_D, _D, _D, _I, _D, ctypes.c_double]),
        "pn1_set_c": (_I, [ctypes.c_double]),
        "pn1_set_mode": (_I, [_I]),
        "pn1_create": (_I, [ctypes.POINTER(P), _I, _I]),
        "pn1_destroy": (None, [P]),
        "pn1_set_params": (_I, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I]),
        "pn1_upload": (_I, [P, _D, _D, _D, _I]),
        "pn1_download": (_I, [P, _D, _D]),
        "pn1_force": (_I, [P, _D]),
        "pn1_force_range": (_I, [P, _I, _I, _D]),
        "pn1_energy": (_I, [P, ctypes.POINTER(Energy)]),
        "pn1_step": (_I, [P, _I]),
        "pn1_sync": (_I, [P]),
        "pn1_last_error": (ctypes.c_char_p, [P]),
        "pn1_set_tuning": (_I, [P, _I, _I, _I]),
        "pn1_get_tuning": (_I, [P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "pn1_profile_enable": (_I, [P, _I]),
        "pn1_profile_read": (_I, [P, ctypes.POINTER(_I), _D]),
        "pn1_autotune": (_I, [P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _D]),
        "pn1_force_noalloc": (_I, [ctypes.c_void_p, ctypes.c_void_p, _I, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
        "pn1_ens_autotune": (_I, [P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _D]),
        "pn1_force_vjp": (_I, [P, _D, _D, _D, _D]),
        "pn1_adjoint_forward": (_I, [P, _I, _I]),
        "pn1_adjoint_backward": (_I, [P, _D, _D, _D, _D, _D]),
        "pn1_ens_create": (_I, [ctypes.POINTER(P), _I, _I, _I]),
        "pn1_ens_destroy": (None, [P]),
        "pn1_ens_set_params": (_I, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I]),
        "pn1_ens_upload": (_I, [P, _I, ctypes.POINTER(_I), _D, _D, _D]),
        "pn1_ens_download": (_I, [P, _D, _D]),
        "pn1_ens_force": (_I, [P, _D]),
        "pn1_ens_energy": (_I, [P, _D]),
        "pn1_ens_step": (_I, [P, _I]),
        "pn1_ens_sync": (_I, [P]),
        "pn1_ens_last_error": (ctypes.c_char_p, [P]),
        "pn1_ens_set_tuning": (_I, [P, _I, _I, _I]),
        "pn1_ens_get_tuning": (_I, [P, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                    ctypes.POINTER(_I)]),
        "pn1_ens_profile_enable": (_I, [P, _I]),
        "pn1_ens_profile_read": (_I, [P, ctypes.POINTER(_I), _D]),
        "pn1_ens_force_vjp": (_I, [P, _D, _D, _D, _D]),
        "pn1_ens_adjoint_forward": (_I, [P, _I, _I]),
        "pn1_ens_adjoint_backward": (_I, [P, _D, _D, _D, _D, _D]),
        "pn1_check_input": (_I, [_D, _D, _D, _I, ctypes.c_double, ctypes.c_char_p, _I]),
        "pn1_ic_uniform_sphere": (_I, [_I, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, _D, _D, _D]),
        "pn1_ic_plummer": (_I, [_I, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double,
                                _D, _D, _D]),
        "pn1_read_dat": (_I, [ctypes.c_char_p, _I, _D, _D]),
        "pn1_write_snapshot": (_I, [ctypes.c_char_p, _I, _D, _D, _I]),
        "pn1_write_checkpoint": (_I, [ctypes.c_char_p, _I, ctypes.c_double, ctypes.c_longlong,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _D, _D, _D]),
        "pn1_read_checkpoint": (_I, [ctypes.c_char_p, ctypes.POINTER(_I), _D,
                                     ctypes.POINTER(ctypes.c_longlong), _D, _D, _D,
                                     _D, _D, _D, _I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a):
    return a.ctypes.data_as(_D)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


# ---------------------------------------------------------------- context
class Context:
    """A device-resident simulation state (pn1_ctx)."""

    def __init__(self, capacity, device=0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.pn1_create(ctypes.byref(h), int(capacity), int(device))
        if rc != PN1_OK:
            raise Pn1Error(rc, f"pn1_create(capacity={capacity}, device={device})")
        self._L, self._h, self.capacity, self.n = L, h, int(capacity), 0

    def _chk(self, rc, what):
        if rc != PN1_OK:
            raise Pn1Error(rc, f"{what}: {self._L.pn1_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self._L.pn1_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, eps, c=1.0, dt=2.0**-10, mode=FP64):
        self._chk(self._L.pn1_set_params(self._h, eps, c, dt, int(mode)), "pn1_set_params")

    def set_tuning(self, jsplit=0, tile=0, variant=-1):
        self._chk(self._L.pn1_set_tuning(self._h, int(jsplit), int(tile), int(variant)),
                  "pn1_set_tuning")

    def get_tuning(self):
        s, t, v = _I(), _I(), _I()
        self._chk(self._L.pn1_get_tuning(self._h, ctypes.byref(s), ctypes.byref(t),
                                         ctypes.byref(v)), "pn1_get_tuning")
        return s.value, t.value, v.value

    def autotune(self, reps=3):
        """(jsplit, variant, pair ms) of the fastest shape, now in use."""
        s, v, ms = _I(), _I(), ctypes.c_double()
        self._chk(self._L.pn1_autotune(self._h, int(reps), ctypes.byref(s), ctypes.byref(v),
                                       ctypes.byref(ms)), "pn1_autotune")
        return s.value, v.value, ms.value

    def upload(self, x, v, m):
        m = _f64(m)
        n = m.shape[0]
        x, v = _f64(x, (n, 3)), _f64(v, (n, 3))
        self._chk(self._L.pn1_upload(self._h, ptr(x), ptr(v), ptr(m), n), "pn1_upload")
        self.n = n

    def download(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        self._chk(self._L.pn1_download(self._h, ptr(x), ptr(v)), "pn1_download")
        return x, v

    def force(self):
        a = np.empty((self.n, 3))
        self._chk(self._L.pn1_force(self._h, ptr(a)), "pn1_force")
        return a

    def force_range(self, i0, i1):
        """Accelerations of stars i0 .. i1 - 1 only (one i shard; 256-star
        block boundaries), [i1 - i0, 3]."""
        a = np.empty((max(int(i1) - int(i0), 0), 3))
        self._chk(self._L.pn1_force_range(self._h, int(i0), int(i1), ptr(a)), "pn1_force_range")
        return a

    def energy(self):
        e = Energy()
        self._chk(self._L.pn1_energy(self._h, ctypes.byref(e)), "pn1_energy")
        return e.as_array()

    def step(self, steps):
        self._chk(self._L.pn1_step(self._h, int(steps)), "pn1_step")

    def sync(self):
        self._chk(self._L.pn1_sync(self._h), "pn1_sync")

    def profile_enable(self, on=True):
        self._chk(self._L.pn1_profile_enable(self._h, 1 if on else 0), "pn1_profile_enable")

    def profile_read(self):
        k, ms = _I(), ctypes.c_double()
        self._chk(self._L.pn1_profile_read(self._h, ctypes.byref(k), ctypes.byref(ms)),
                  "pn1_profile_read")
        return k.value, ms.value

    # ---- adjoint
    def force_vjp(self, u):
        """(gx, gv, gm0): gradient of sum_p u_p . a_p at the resident state."""
        u = _f64(u, (self.n, 3))
        gx = np.empty((self.n, 3))
        gv = np.empty((self.n, 3))
        gm0 = ctypes.c_double()
        self._chk(self._L.pn1_force_vjp(self._h, ptr(u), ptr(gx), ptr(gv), ctypes.byref(gm0)),
                  "pn1_force_vjp")
        return gx, gv, gm0.value

    def adjoint_forward(self, steps, interval=16):
        self._chk(self._L.pn1_adjoint_forward(self._h, int(steps), int(interval)),
                  "pn1_adjoint_forward")

    def adjoint_backward(self, gx, gv):
        """(gx0, gv0, gm0) for the final-state gradient (gx, gv)."""
        gx, gv = _f64(gx, (self.n, 3)), _f64(gv, (self.n, 3))
        gx0 = np.empty((self.n, 3))
        gv0 = np.empty((self.n, 3))
        gm0 = ctypes.c_double()
        self._chk(self._L.pn1_adjoint_backward(self._h, ptr(gx), ptr(gv), ptr(gx0), ptr(gv0),
                                               ctypes.byref(gm0)), "pn1_adjoint_backward")
        return gx0, gv0, gm0.value


# ---------------------------------------------------------------- ensemble
class Ensemble:
    """Many independent clusters advanced together (pn1_ens).

    Clusters are given as a list of (x, v, m) states, each with its BH
    first; arrays returned by `force` / `download` are lists in the same
    order, and `energy` returns an array [nclusters, 5] of
    (kin, pot, pot_pn, pot_pn2, total)."""

    def __init__(self, capacity, max_clusters, device=0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.pn1_ens_create(ctypes.byref(h), int(capacity), int(max_clusters), int(device))
        if rc != PN1_OK:
            raise Pn1Error(rc, f"pn1_ens_create(capacity={capacity}, "
                               f"max_clusters={max_clusters}, device={device})")
        self._L, self._h = L, h
        self.capacity, self.max_clusters = int(capacity), int(max_clusters)
        self.offsets = np.zeros(1, dtype=np.int32)

    def _chk(self, rc, what):
        if rc != PN1_OK:
            raise Pn1Error(rc, f"{what}: {self._L.pn1_ens_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self._L.pn1_ens_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return int(self.offsets[-1])

    def split(self, a):
        o = self.offsets
        return [a[o[c]:o[c + 1]] for c in range(len(o) - 1)]

    def set_params(self, eps, c=1.0, dt=2.0**-10, mode=FP64):
        self._chk(self._L.pn1_ens_set_params(self._h, eps, c, dt, int(mode)),
                  "pn1_ens_set_params")

    def set_tuning(self, chunk=0, tile=0, variant=-1):
        self._chk(self._L.pn1_ens_set_tuning(self._h, int(chunk), int(tile), int(variant)),
                  "pn1_ens_set_tuning")

    def get_tuning(self):
        s, t, v = _I(), _I(), _I()
        self._chk(self._L.pn1_ens_get_tuning(self._h, ctypes.byref(s), ctypes.byref(t),
                                             ctypes.byref(v)), "pn1_ens_get_tuning")
        return s.value, t.value, v.value

    def autotune(self, reps=3):
        """(chunk, variant, pair ms) of the fastest shape, now in use."""
        s, v, ms = _I(), _I(), ctypes.c_double()
        self._chk(self._L.pn1_ens_autotune(self._h, int(reps), ctypes.byref(s), ctypes.byref(v),
                                           ctypes.byref(ms)), "pn1_ens_autotune")
        return s.value, v.value, ms.value

    def upload(self, clusters):
        sizes = [len(m) for _, _, m in clusters]
        off = np.zeros(len(sizes) + 1, dtype=np.int32)
        off[1:] = np.cumsum(sizes)
        def cat(k, shape):
            parts = [np.asarray(c[k], dtype=np.float64).reshape(shape) for c in clusters]
            return _f64(np.concatenate(parts) if parts else np.zeros((1,) + shape[1:]))
        x, v, m = cat(0, (-1, 3)), cat(1, (-1, 3)), cat(2, (-1,))
        self._chk(self._L.pn1_ens_upload(self._h, len(sizes),
                                         off.ctypes.data_as(ctypes.POINTER(_I)),
                                         ptr(x), ptr(v), ptr(m)), "pn1_ens_upload")
        self.offsets = off

    def download(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        self._chk(self._L.pn1_ens_download(self._h, ptr(x), ptr(v)), "pn1_ens_download")
        return self.split(x), self.split(v)

    def force(self):
        a = np.empty((self.n, 3))
        self._chk(self._L.pn1_ens_force(self._h, ptr(a)), "pn1_ens_force")
        return self.split(a)

    def energy(self):
        e = np.empty((len(self.offsets) - 1, 5))
        self._chk(self._L.pn1_ens_energy(self._h, ptr(e)), "pn1_ens_energy")
        return e

    def step(self, steps):
        self._chk(self._L.pn1_ens_step(self._h, int(steps)), "pn1_ens_step")

    def sync(self):
        self._chk(self._L.pn1_ens_sync(self._h), "pn1_ens_sync")

    def profile_enable(self, on=True):
        self._chk(self._L.pn1_ens_profile_enable(self._h, 1 if on else 0),
                  "pn1_ens_profile_enable")

    def profile_read(self):
        k, ms = _I(), ctypes.c_double()
        self._chk(self._L.pn1_ens_profile_read(self._h, ctypes.byref(k), ctypes.byref(ms)),
                  "pn1_ens_profile_read")
        return k.value, ms.value

    # ---- adjoint
    def _cat3(self, parts):
        return _f64(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 3)
                                    for p in parts]), (self.n, 3))

    def force_vjp(self, us):
        """Per cluster (gx, gv) lists and the m0 derivatives, for the
        per-cluster cotangents `us`."""
        u = self._cat3(us)
        gx = np.empty((self.n, 3))
        gv = np.empty((self.n, 3))
        gm0 = np.empty(len(self.offsets) - 1)
        self._chk(self._L.pn1_ens_force_vjp(self._h, ptr(u), ptr(gx), ptr(gv), ptr(gm0)),
                  "pn1_ens_force_vjp")
        return self.split(gx), self.split(gv), gm0

    def adjoint_forward(self, steps, interval=16):
        self._chk(self._L.pn1_ens_adjoint_forward(self._h, int(steps), int(interval)),
                  "pn1_ens_adjoint_forward")

    def adjoint_backward(self, gxs, gvs):
        gx, gv = self._cat3(gxs), self._cat3(gvs)
        gx0 = np.empty((self.n, 3))
        gv0 = np.empty((self.n, 3))
        gm0 = np.empty(len(self.offsets) - 1)
        self._chk(self._L.pn1_ens_adjoint_backward(self._h, ptr(gx), ptr(gv), ptr(gx0), ptr(gv0),
                                                   ptr(gm0)), "pn1_ens_adjoint_backward")
        return self.split(gx0), self.split(gv0), gm0


# ---------------------------------------------------------------- no-allocation force
NOALLOC_MAX_N = 16384


def noalloc_split(n):
    """The j split pn1_force_noalloc uses for n particles."""
    nst, s = n - 1, 32
    while s > 1 and (nst + s - 1) // s < 64:
        s //= 2
    return s


def force_noalloc(st, m, n, eps, c, acc, stream=0):
    """pn1_force_noalloc on device addresses (ints): st 6 x n doubles, m n
    doubles, acc 3 x n doubles out; enqueued on `stream` (a hipStream_t
    address, 0 for the null stream)."""
    rc = load().pn1_force_noalloc(ctypes.c_void_p(st), ctypes.c_void_p(m), int(n), float(eps),
                                  float(c), ctypes.c_void_p(acc), ctypes.c_void_p(stream))
    if rc != PN1_OK:
        raise Pn1Error(rc, f"pn1_force_noalloc(n={n}, eps={eps}, c={c}) failed")


# ---------------------------------------------------------------- legacy shapes
def legacy_force(x, v, m, eps):
    """force() with the reference's prototype; returns a [n][3]."""
    L = load()
    m = _f64(m)
    n = m.shape[0]
    x, v = _f64(x, (n, 3)), _f64(v, (n, 3))
    a = np.zeros((n, 3))
    pot = np.zeros(n)
    L.force(ptr(x), ptr(v), ptr(m), eps, ptr(a), ptr(pot), n)
    return a


def set_c(c):
    if load().pn1_set_c(float(c)) != PN1_OK:
        raise Pn1Error(PN1_EINVAL, f"pn1_set_c({c})")


def set_mode(mode):
    if load().pn1_set_mode(int(mode)) != PN1_OK:
        raise Pn1Error(PN1_EINVAL, f"pn1_set_mode({mode})")


# ---------------------------------------------------------------- host helpers
def check_input(x, v, m, eps):
    L = load()
    m = _f64(m)
    n = m.shape[0]
    x, v = _f64(x, (n, 3)), _f64(v, (n, 3))
    buf = ctypes.create_string_buffer(256)
    rc = L.pn1_check_input(ptr(x), ptr(v), ptr(m), n, eps, buf, 256)
    return rc, buf.value.decode()


def ic_uniform_sphere(n_stars, seed, R=1.0, Q=0.5, m_bh=0.01, eps=1e-3):
    L = load()
    n = n_stars + 1
    x, v, m = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    rc = L.pn1_ic_uniform_sphere(n_stars, seed, R, Q, m_bh, eps, ptr(x), ptr(v), ptr(m))
    if rc != PN1_OK:
        raise Pn1Error(rc, "pn1_ic_uniform_sphere")
    return x, v, m


def ic_plummer(n_stars, seed, a=1.0, m_bh=0.05):
    L = load()
    n = n_stars + 1
    x, v, m = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    rc = L.pn1_ic_plummer(n_stars, seed, a, m_bh, ptr(x), ptr(v), ptr(m))
    if rc != PN1_OK:
        raise Pn1Error(rc, "pn1_ic_plummer")
    return x, v, m


def read_dat(path, n):
    L = load()
    x, v = np.zeros((n, 3)), np.zeros((n, 3))
    rc = L.pn1_read_dat(os.fsencode(path), n, ptr(x), ptr(v))
    if rc != PN1_OK:
        raise Pn1Error(rc, f"pn1_read_dat({path}, n={n})")
    return x, v


def write_snapshot(path, x, v, append=False):
    L = load()
    x, v = _f64(x), _f64(v)
    rc = L.pn1_write_snapshot(os.fsencode(path), 1 if append else 0, ptr(x), ptr(v), x.shape[0])
    if rc != PN1_OK:
        raise Pn1Error(rc, f"pn1_write_snapshot({path})")


def write_checkpoint(path, x, v, m, t=0.0, step=0, eps=1e-3, c=1.0, dt=2.0**-10):
    L = load()
    x, v, m = _f64(x), _f64(v), _f64(m)
    rc = L.pn1_write_checkpoint(os.fsencode(path), m.shape[0], t, step, eps, c, dt,
                                ptr(x), ptr(v), ptr(m))
    if rc != PN1_OK:
        raise Pn1Error(rc, f"pn1_write_checkpoint({path})")


def read_checkpoint(path):
    L = load()
    n = _I()
    t, eps, c, dt = (ctypes.c_double() for _ in range(4))
    step = ctypes.c_longlong()
    null = ctypes.cast(None, _D)
    rc = L.pn1_read_checkpoint(os.fsencode(path), ctypes.byref(n), ctypes.byref(t),
                               ctypes.byref(step), ctypes.byref(eps), ctypes.byref(c),
                               ctypes.byref(dt), null, null, null, 0)
    if rc != PN1_OK:
        raise Pn1Error(rc, f"pn1_read_checkpoint({path})")
    N = n.value
    x, v, m = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros(N)
    rc = L.pn1_read_checkpoint(os.fsencode(path), ctypes.byref(n), None, None, None, None, None,
                               ptr(x), ptr(v), ptr(m), N)
    if rc != PN1_OK:
        raise Pn1Error(rc, f"pn1_read_checkpoint({path})")
    return dict(x=x, v=v, m=m, t=t.value, step=step.value, eps=eps.value, c=c.value, dt=dt.value)
