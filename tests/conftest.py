import os
import sys

import numpy as np
import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(TESTS)
GOLDEN = os.path.join(TESTS, "golden")
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


FORCE_FIXTURES = ["rand_n2", "rand_n3", "rand_n17", "rand_n65", "rand_n257", "close_n65",
                  "dat_n65", "c1_n1025"]


def load_fixture(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(params=FORCE_FIXTURES)
def force_fixture(request):
    d = load_fixture(request.param)
    d["name"] = request.param
    return d


def rel_err(a, ref):
    """Per-particle ||a_i - ref_i|| / ||ref_i||."""
    num = np.linalg.norm(a - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    return num / np.where(den > 0, den, 1.0)


def c_values(d):
    return sorted(int(k[3:]) for k in d if k.startswith("a_c") and k != "a_c1")


def energy_atol(d, ref):
    """Absolute tolerance for energy terms.  The reference adds each self
    pair (w = 1/eps) and removes it again (Limit-sticky9.c:472, 548-549), so
    its sums carry rounding of order ulp(sum m_i^2 / eps) whatever the size
    of the result; beyond that, 1e-13 of the largest term."""
    m, eps = d["m"], float(d["eps"])
    self_scale = float((m * m).sum()) / eps
    return 4e-16 * self_scale * 4 + 1e-13 * np.abs(ref).max()
