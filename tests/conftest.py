import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: full-size case (tens of seconds)")


@pytest.fixture(scope="session")
def golden():
    with np.load(os.path.join(GOLDEN, "ref_small.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_c1():
    path = os.path.join(GOLDEN, "ref_c1.npz")
    if not os.path.exists(path):
        pytest.skip("ref_c1.npz fixture not generated")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def nb():
    """The HIP library.  On the GPU box a missing build is a failure, not
    a skip: GPU tests must run the native path."""
    import nekbone_amd
    nekbone_amd.lib()
    nekbone_amd.init(0)
    return nekbone_amd


def splitmix_uniform(n, seed):
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return 2.0 * u - 1.0


SMALL_CASES = [((2, 2, 2), 8), ((3, 2, 2), 10), ((1, 1, 1), 5), ((2, 3, 1), 12), ((4, 1, 1), 2),
               ((2, 1, 3), 6)]


def tag_of(nel, nx1):
    return "b%d%d%d_n%d" % (nel[0], nel[1], nel[2], nx1)
